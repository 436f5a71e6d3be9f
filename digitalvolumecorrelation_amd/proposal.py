"""Learned start-warp proposal (dvc_net_* / dvc_search_propose in include/dvc_amd.h).

The reference starts every point's Gauss-Newton search from the mean converged
translation of its earlier neighbours (Core/Search.cpp:277-377), which chains all
points of an ROI into one serial order.  Here a small 3D convolutional network,
trained on the GPU by the library's own kernels, proposes all 12 warp parameters of a
point from the 32^3 reference/target crops around it.  Every point can then be
searched at once (START_GIVEN), and a judge pass from the reference's own start
guess keeps the proposal's result only where it is no worse (START_JUDGED).

Device buffers are the context's own allocations (DeviceArray over dvc_device_alloc);
every kernel is the library's, queued on the context's stream, and DeviceArray copies
wait for that stream.
"""
from __future__ import annotations

import ctypes as C
import time
from dataclasses import dataclass
from pathlib import Path

import numpy as np

from . import _lib
from ._lib import Context, DeviceArray, DvcError

SIDE = 32
CROP_ELEMS = SIDE ** 3 * 2                     # BF16 values per crop (2 channels)
SCALE = np.array([8.0] * 3 + [0.05] * 3 + [0.02] * 6)
WEIGHTS = Path(__file__).resolve().parent / "weights" / "proposal_c16_32_64_128.npz"


def num_params() -> int:
    return int(_lib.load().dvc_net_num_params())


def init_params(seed: int = 0) -> np.ndarray:
    """He-normal convolution weights, small head weights, zero biases; the layout of
    dvc_net_params (per convolution [cout][27*cin] then [cout]; head [12][128], [12])."""
    rng = np.random.default_rng(seed)
    parts = []
    for cin, cout in zip((2, 16, 32, 64), (16, 32, 64, 128)):
        k = 27 * cin
        parts += [rng.normal(0.0, np.sqrt(2.0 / k), cout * k), np.zeros(cout)]
    parts += [rng.normal(0.0, 0.1 / np.sqrt(128), 12 * 128), np.zeros(12)]
    flat = np.concatenate(parts).astype(np.float32)
    assert flat.size == num_params()
    return flat


class ProposalNet:
    """dvc_net on one context: max_batch crops per forward; training=True adds the
    gradient and Adam buffers."""

    def __init__(self, ctx: Context, max_batch: int = 256, training: bool = False, params=None):
        self.ctx = ctx
        self.lib = ctx.lib
        self.max_batch = max_batch
        h = C.c_void_p()
        rc = self.lib.dvc_net_create(ctx.h, max_batch, int(training), C.byref(h))
        if rc != 0:
            raise DvcError(rc, "dvc_net_create failed")
        self.h = h
        ctx._adopt(self)
        self.set_params(init_params(0) if params is None else params)

    def close(self):
        # a network must be destroyed before its context (dvc_net_destroy waits on the
        # context's stream); Context.close() closes its networks first
        if getattr(self, "h", None) and getattr(self.ctx, "h", None):
            self.lib.dvc_net_destroy(self.h)
        self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, rc):
        if rc != 0:
            raise DvcError(rc, self.lib.dvc_net_last_error(self.h).decode())

    # ---- parameters
    def set_params(self, flat):
        f = np.ascontiguousarray(flat, dtype=np.float32)
        self._check(self.lib.dvc_net_params(self.h, f.ctypes.data_as(C.POINTER(C.c_float)), f.size, 1))

    def params(self) -> np.ndarray:
        f = np.zeros(num_params(), np.float32)
        self._check(self.lib.dvc_net_params(self.h, f.ctypes.data_as(C.POINTER(C.c_float)), f.size, 0))
        return f

    def save(self, path, **meta):
        path = Path(path)
        path.parent.mkdir(parents=True, exist_ok=True)
        np.savez(path, params=self.params(), **{k: np.asarray(v) for k, v in meta.items()})

    @staticmethod
    def load_params(path=WEIGHTS) -> np.ndarray:
        with np.load(path, allow_pickle=False) as z:
            return np.array(z["params"], dtype=np.float32)

    # ---- kernels
    def crop(self, points_dev: int, count: int, out_dev: int | None = None, idx_dev: int | None = None,
             slot_dev: int | None = None, flip_dev: int | None = None):
        """Crops around device points (float64 [n, 3]) into out_dev (BF16 [slots, 32^3, 2]),
        crop i at slot slot_dev[i] (default i), mirrored along the axes of flip_dev[i]."""
        v = lambda x: C.c_void_p(x) if x else None  # noqa: E731
        self._check(self.lib.dvc_net_crop(self.h, C.c_void_p(points_dev), v(idx_dev), count, v(out_dev), v(slot_dev),
                                          v(flip_dev)))

    def forward(self, in_dev: int | None, B: int, par_dev: int | None = None, y_dev: int | None = None):
        self._check(self.lib.dvc_net_forward(self.h, C.c_void_p(in_dev) if in_dev else None, B,
                                             C.c_void_p(par_dev) if par_dev else None,
                                             C.c_void_p(y_dev) if y_dev else None))

    def gradients(self, in_dev: int | None, label_dev: int, B: int):
        g = np.zeros(num_params(), np.float32)
        loss = C.c_float()
        self._check(self.lib.dvc_net_gradients(self.h, C.c_void_p(in_dev) if in_dev else None, C.c_void_p(label_dev),
                                               B, g.ctypes.data_as(C.POINTER(C.c_float)), C.byref(loss)))
        return float(loss.value), g

    def train_step(self, in_dev: int | None, label_dev: int, B: int, lr: float, want_loss=False):
        loss = C.c_float()
        self._check(self.lib.dvc_net_train_step(self.h, C.c_void_p(in_dev) if in_dev else None, C.c_void_p(label_dev),
                                                B, float(lr), C.byref(loss) if want_loss else None))
        return float(loss.value) if want_loss else None

    def propose(self):
        """Start warps of every point of the context's prepared given-start / judged
        search (dvc_search_propose)."""
        self._check(self.lib.dvc_search_propose(self.ctx.h, self.h))

    # ---- convenience on host arrays (tests, small runs)
    def predict_points(self, points: np.ndarray) -> np.ndarray:
        """Warp proposals [n, 12] for host points, from the context's volume pair."""
        pts = DeviceArray.from_host(self.ctx, np.ascontiguousarray(points, dtype=np.float64).reshape(-1, 3))
        out = DeviceArray(self.ctx, (len(points), 12), np.float64)
        crops = DeviceArray(self.ctx, (self.max_batch, CROP_ELEMS), np.uint16)
        for i0 in range(0, len(points), self.max_batch):
            b = min(self.max_batch, len(points) - i0)
            self.crop(pts.row(i0), b, crops.ptr)
            self.forward(crops.ptr, b, out.row(i0))
        return out.download()


# ---------------------------------------------------------------- training data

def _rotation(axis, deg):
    a = np.asarray(axis, dtype=np.float64)
    a = a / np.linalg.norm(a)
    t = np.deg2rad(deg)
    K = np.array([[0, -a[2], a[1]], [a[2], 0, -a[0]], [-a[1], a[0], 0]])
    return np.eye(3) + np.sin(t) * K + (1 - np.cos(t)) * K @ K


@dataclass
class TiledPair:
    """A training pair whose target is the reference warped tile by tile, each tile
    under its own affine map phi_t(X) = c_t + t_t + F_t (X - c_t) about the tile centre:
    thousands of independent deformations per volume pair."""
    ref: np.ndarray
    maps: np.ndarray          # [ntiles, 12] inverse maps (dvc_synth_warp_tiles)
    points: np.ndarray        # [n, 3]
    disp: np.ndarray          # [n, 3] imposed displacement phi_t(X) - X at each point
    tile_of: np.ndarray       # [n] tile index


def reference_volume(seed: int, size=384):
    """Smoothed-noise uint8 reference (sigma 1.5-2.5 drawn from the seed)."""
    from . import synth
    rng = np.random.default_rng(30_000 + seed)
    return synth.smoothed_noise((size,) * 3, rng.uniform(1.5, 2.5), np.uint8, seed=20_000 + seed)


def tiled_pair(seed: int, size=384, tile=64, trans=8.0, rot_deg=3.0, strain=0.02, spacing=3.0, per_axis=5, ref=None):
    """Reference volume (smoothed noise, sigma 1.5-2.5, or `ref`) and per-tile affine
    maps drawn from `seed`: translation up to `trans` voxels per axis, rotation up to
    `rot_deg` about a random axis through the tile centre, symmetric strain up to
    `strain`; per_axis^3 points around each tile centre at `spacing`."""
    rng = np.random.default_rng(seed)
    if ref is None:
        ref = reference_volume(seed, size)
    size = ref.shape[0]
    nt = (size + tile - 1) // tile
    maps, pts, disp, tix = [], [], [], []
    g = (np.arange(per_axis) - (per_axis - 1) / 2.0) * spacing
    gz, gy, gx = np.meshgrid(g, g, g, indexing="ij")
    local = np.stack([gx.ravel(), gy.ravel(), gz.ravel()], axis=1)
    for tz in range(nt):
        for ty in range(nt):
            for tx in range(nt):
                c = np.array([tx, ty, tz], dtype=np.float64) * tile + (tile - 1) / 2.0
                t = rng.uniform(-trans, trans, 3)
                E = rng.uniform(-strain, strain, (3, 3))
                F = _rotation(rng.normal(size=3), rng.uniform(-rot_deg, rot_deg)) @ (np.eye(3) + 0.5 * (E + E.T))
                Fi = np.linalg.inv(F)
                maps.append(np.concatenate([Fi.ravel(), c - Fi @ (c + t)]))
                X = c + local
                pts.append(X)
                disp.append(t + (X - c) @ (F - np.eye(3)).T)
                tix.append(np.full(len(X), len(maps) - 1))
    return TiledPair(ref, np.array(maps), np.concatenate(pts), np.concatenate(disp), np.concatenate(tix))


LABEL_PARAMS = dict(ndof=12, obj="znssd", interp="tricubic", geom="cube", subvol_size=21, subvol_npts=1331,
                    disp_max=12, basin_radius=0.0)


def device_reference(ctx: Context, seed: int, size=384):
    """A smoothed-noise uint8 reference made in slot 0 on the device (dvc_synth_noise,
    sigma 1.5-2.5 drawn from the seed)."""
    rng = np.random.default_rng(40_000 + seed)
    ctx.synth_noise(0, (size,) * 3, np.uint8, float(rng.uniform(1.5, 2.5)), 50_000 + seed)


def label_pair(ctx: Context, tp: TiledPair, tile=64, params=None, on_device=False):
    """Upload the pair (target made on the device) and label its points with the plain
    search's converged 12-DOF warps, each point started from its tile's imposed
    translation (DVC_START_GIVEN).  Keeps points that converged (good status, fewer than
    20 Gauss-Newton evaluations) within half a voxel of the imposed displacement.
    Leaves the pair in the context; returns (points, labels [n, 12])."""
    p = dict(LABEL_PARAMS if params is None else params)
    if not on_device:
        ctx.upload(0, tp.ref)
    ctx.warp_tiles(tile, tp.maps)
    order, neigh = _lib.knn_order_neighbors(tp.points, k=8)
    starts = np.zeros((len(tp.points), 12))
    starts[:, :3] = tp.disp
    ctx.set_start_mode(_lib.START_GIVEN)
    try:
        ctx.prepare(_lib.make_params(**p), tp.points, order, neigh)
        ctx.set_starts(starts)
        ctx.execute()
        res = ctx.fetch()
    finally:
        ctx.set_start_mode(_lib.START_CHAIN)
    ok = (res["status"] == 0) & (res["iters"] < 20) & (np.abs(res["par"][:, :3] - tp.disp).max(axis=1) < 0.5)
    return tp.points[ok], res["par"][ok].copy()


def flip_labels(labels: np.ndarray, flips: np.ndarray) -> np.ndarray:
    """Warp parameters of mirrored crop pairs.  Mirroring axes with signs s = diag(D)
    maps the warp's rotation and strain to D A D and D S D: u_a -> s_a u_a; the angles
    phi, theta, psi (rotation in the xy, xz, yz planes, SearchParams.cpp:53-107) and the
    shears exy, eyz, exz take the product of their two axes' signs; normal strains keep
    theirs."""
    sx = np.where(flips & 1, -1.0, 1.0)
    sy = np.where(flips & 2, -1.0, 1.0)
    sz = np.where(flips & 4, -1.0, 1.0)
    one = np.ones_like(sx)
    f = np.stack([sx, sy, sz, sx * sy, sx * sz, sy * sz, one, one, one, sx * sy, sy * sz, sx * sz], axis=1)
    return labels * f


@dataclass
class TrainSet:
    crops: DeviceArray     # uint16 [N, 65536] (BF16 crop pairs), device
    labels: DeviceArray    # float32 [N, 12], device
    labels_host: np.ndarray
    n: int


def build_train_set(ctx: Context, net: ProposalNet, seeds, size=384, tile=64, log=print, shuffle_seed=0,
                    maps_per_ref=1, per_axis=5, spacing=3.0, flip=True, device_textures=False, trans=8.0):
    """Crops and plain-search labels of tiled pairs, resident in HBM, in a random order
    over all pairs (every contiguous batch mixes them).  Each reference volume serves
    `maps_per_ref` pairs with different tile maps, so a texture seen in training comes
    with several deformations and cannot stand for one."""
    pairs, labels = [], []
    ref = None
    for i, s in enumerate(seeds):
        t0 = time.time()
        if device_textures:
            device_reference(ctx, int(s), size)
            ref = np.zeros((size, size, size), np.uint8)     # shape only; the texture is in slot 0
        elif i % maps_per_ref == 0:
            ref = reference_volume(int(s), size)
        tp = tiled_pair(int(s), size=size, tile=tile, per_axis=per_axis, spacing=spacing, ref=ref, trans=trans)
        pts, lab = label_pair(ctx, tp, tile, on_device=device_textures)
        if device_textures:
            # crop now, while the pair is in the context (slots hold one pair at a time)
            tp.ref = None
        pairs.append((tp, pts, (int(s) if device_textures else None)))
        labels.append(lab)
        log(f"training pair {s}: {len(pts)} labelled points of {len(tp.points)} in {len(tp.maps)} tiles "
            f"({time.time() - t0:.1f} s)")
    n = sum(len(l) for l in labels)
    rng = np.random.default_rng(shuffle_seed)
    perm = rng.permutation(n).astype(np.int32)
    flips = rng.integers(0, 8, n).astype(np.int32) if flip else np.zeros(n, np.int32)
    crops = DeviceArray(ctx, (n, CROP_ELEMS), np.uint16)
    lab_all = flip_labels(np.concatenate(labels), flips).astype(np.float32)
    lab_store = np.empty_like(lab_all)
    lab_store[perm] = lab_all                 # sample i is stored at slot perm[i]
    off = 0
    for tp, pts, dev_seed in pairs:
        if dev_seed is not None:
            device_reference(ctx, dev_seed, size)
        else:
            ctx.upload(0, tp.ref)
        ctx.warp_tiles(tile, tp.maps)
        dp = DeviceArray.from_host(ctx, pts)
        slots = DeviceArray.from_host(ctx, perm[off:off + len(pts)])
        fl = DeviceArray.from_host(ctx, flips[off:off + len(pts)])
        net.crop(dp.ptr, len(pts), crops.ptr, slot_dev=slots.ptr, flip_dev=fl.ptr)
        ctx.synchronize()
        off += len(pts)
    return TrainSet(crops, DeviceArray.from_host(ctx, lab_store), lab_store, n)


def train(net: ProposalNet, ts: TrainSet, held: TrainSet, epochs=4, batch=256, lr=1e-3, lr_final=1e-4, log=print):
    """Adam over contiguous batches of the training set in a random order per epoch,
    cosine-decayed learning rate; the loss of `held` (pairs never trained on) is
    reported per epoch.  Returns the per-epoch (train, held-out) losses."""
    nb = ts.n // batch
    steps = nb * epochs
    rng = np.random.default_rng(1)
    hist = []
    k = 0
    best, best_params = None, None
    for ep in range(epochs):
        t0 = time.time()
        run = []
        for bi in rng.permutation(nb):
            f = k / max(1, steps - 1)
            rate = lr_final + 0.5 * (lr - lr_final) * (1 + np.cos(np.pi * f))
            off = int(bi) * batch
            want = (k % 50) == 0
            l = net.train_step(ts.crops.row(off), ts.labels.row(off), batch, rate, want_loss=want)
            if want:
                run.append(l)
            k += 1
        hl = evaluate(net, held, 0, held.n, batch)
        hist.append((float(np.mean(run)) if run else float("nan"), hl))
        if best is None or hl < best:
            best, best_params = hl, net.params()
        log(f"epoch {ep + 1}/{epochs}: train loss {hist[-1][0]:.4f}, held-out loss {hl:.4f} "
            f"({time.time() - t0:.1f} s, {nb} steps of {batch})")
    # keep the epoch with the lowest held-out loss (early stopping)
    net.set_params(best_params)
    return hist


def evaluate(net: ProposalNet, ts: TrainSet, i0: int, i1: int, batch=256, per_output=False):
    """Mean loss of samples i0..i1 (forward only); per_output: the mean squared error
    of each of the 12 scaled outputs instead."""
    y = DeviceArray(net.ctx, (batch, 12), np.float32)
    tot, cnt = np.zeros(12), 0
    for o in range(i0, i1, batch):
        b = min(batch, i1 - o)
        net.forward(ts.crops.row(o), b, None, y.ptr)
        e = y.download()[:b].astype(np.float64) - ts.labels_host[o:o + b] / SCALE
        tot += (e * e).sum(axis=0)
        cnt += b
    tot /= max(1, cnt)
    return tot if per_output else float(tot.sum())
