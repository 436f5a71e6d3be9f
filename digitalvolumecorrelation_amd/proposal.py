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
        self.set_params(init_params(0) if params is None else params)

    def close(self):
        if getattr(self, "h", None):
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
             slot_dev: int | None = None):
        """Crops around device points (float64 [n, 3]) into out_dev (BF16 [slots, 32^3, 2]),
        crop i at slot slot_dev[i] (default i)."""
        v = lambda x: C.c_void_p(x) if x else None  # noqa: E731
        self._check(self.lib.dvc_net_crop(self.h, C.c_void_p(points_dev), v(idx_dev), count, v(out_dev), v(slot_dev)))

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

@dataclass
class TrainVolume:
    """One synthetic pair with a random smooth deformation: target(x) = ref(psi(x)),
    psi = the inverse of an affine map (translation, rotation about a random axis
    through the centre, symmetric strain) minus a sinusoidal field."""
    ref: np.ndarray
    tar: np.ndarray
    points: np.ndarray
    start_disp: np.ndarray      # displacement at the first point of the order (for rigid_trans)


def _rotation(axis, deg):
    a = np.asarray(axis, dtype=np.float64)
    a = a / np.linalg.norm(a)
    t = np.deg2rad(deg)
    K = np.array([[0, -a[2], a[1]], [a[2], 0, -a[0]], [-a[1], a[0], 0]])
    return np.eye(3) + np.sin(t) * K + (1 - np.cos(t)) * K @ K


def train_volume(seed: int, size: int = 128, spacing: float = 3.0, margin: float = 31.0,
                 trans=8.0, rot_deg=3.0, strain=0.02, amp=1.0, wavelength=64.0) -> TrainVolume:
    from . import synth
    rng = np.random.default_rng(seed)
    sigma = rng.uniform(1.5, 2.5)
    ref = synth.smoothed_noise((size,) * 3, sigma, np.uint8, seed=10_000 + seed)
    c = np.full(3, (size - 1) / 2.0)
    t = rng.uniform(-trans, trans, 3)
    R = _rotation(rng.normal(size=3), rng.uniform(-rot_deg, rot_deg))
    E = rng.uniform(-strain, strain, (3, 3))
    E = 0.5 * (E + E.T)
    F = R @ (np.eye(3) + E)
    Fi = np.linalg.inv(F)
    k = rng.normal(size=(3, 3))
    k = k / np.linalg.norm(k, axis=1, keepdims=True) * (2 * np.pi / wavelength)
    ph = rng.uniform(0, 2 * np.pi, 3)

    def psi(x):
        X = c + (x - c - t) @ Fi.T
        return X - amp * np.sin(x @ k.T + ph)

    tar = synth.warp(ref, psi)
    n = int(np.floor((size - 1 - 2 * margin) / spacing)) + 1
    off = margin + np.arange(n) * spacing
    z, y, x = np.meshgrid(off, off, off, indexing="ij")
    pts = np.stack([x.ravel(), y.ravel(), z.ravel()], axis=1)
    # displacement at the volume centre: x with psi(x) = c, by fixed-point iteration
    xc = c.copy()
    for _ in range(50):
        xc = xc + (c - psi(xc[None])[0])
    return TrainVolume(ref, tar, pts, xc - c)


LABEL_PARAMS = dict(ndof=12, obj="znssd", interp="tricubic", geom="cube", subvol_size=31, subvol_npts=2197,
                    disp_max=12, basin_radius=0.0)


def label_volume(ctx: Context, tv: TrainVolume, params=None):
    """Converged 12-DOF warps of the plain search (the serial chain from the volume
    centre) at the volume's points: (points, labels [n, 12]) of the points whose search
    converged (good status, fewer than 20 Gauss-Newton evaluations)."""
    p = dict(LABEL_PARAMS if params is None else params)
    p["rigid_trans"] = tuple(float(np.rint(v)) for v in tv.start_disp)
    size = tv.ref.shape[0]
    centre = np.full(3, (size - 1) / 2.0)
    order, neigh = _lib.knn_order_neighbors(tv.points, start=centre)
    ctx.upload(0, tv.ref)
    ctx.upload(1, tv.tar)
    res = ctx.search(_lib.make_params(**p), tv.points, order, neigh)
    ok = (res["status"] == 0) & (res["iters"] < 20)
    return tv.points[ok], res["par"][ok].copy()


@dataclass
class TrainSet:
    crops: DeviceArray     # uint16 [N, 65536] (BF16 crop pairs), device
    labels: DeviceArray    # float32 [N, 12], device
    labels_host: np.ndarray
    n: int


def build_train_set(ctx: Context, net: ProposalNet, seeds, size=128, spacing=3.0, log=print, shuffle_seed=0):
    """Crops and plain-search labels of the training volumes, resident in HBM, stored
    in a random order over all volumes (so every contiguous batch mixes them)."""
    vols, labels = [], []
    for s in seeds:
        t0 = time.time()
        tv = train_volume(int(s), size=size, spacing=spacing)
        pts, lab = label_volume(ctx, tv)
        vols.append((tv, pts))
        labels.append(lab)
        log(f"training volume {s}: {len(pts)} labelled points of {len(tv.points)} ({time.time() - t0:.1f} s)")
    n = sum(len(l) for l in labels)
    perm = np.random.default_rng(shuffle_seed).permutation(n).astype(np.int32)
    crops = DeviceArray(ctx, (n, CROP_ELEMS), np.uint16)
    lab_all = np.concatenate(labels).astype(np.float32)
    lab_store = np.empty_like(lab_all)
    lab_store[perm] = lab_all                 # sample i is stored at slot perm[i]
    off = 0
    for tv, pts in vols:
        ctx.upload(0, tv.ref)
        ctx.upload(1, tv.tar)
        dp = DeviceArray.from_host(ctx, pts)
        slots = DeviceArray.from_host(ctx, perm[off:off + len(pts)])
        net.crop(dp.ptr, len(pts), crops.ptr, slot_dev=slots.ptr)
        ctx.synchronize()
        off += len(pts)
    return TrainSet(crops, DeviceArray.from_host(ctx, lab_store), lab_store, n)


def train(net: ProposalNet, ts: TrainSet, epochs=4, batch=256, lr=1e-3, lr_final=1e-4, log=print, holdout=2048):
    """Adam over contiguous batches of the training set in a random order per epoch,
    cosine-decayed learning rate; the last `holdout` samples are held out and their loss
    reported per epoch.  Returns the per-epoch (train, held-out) losses."""
    n_train = ts.n - holdout
    nb = n_train // batch
    steps = nb * epochs
    rng = np.random.default_rng(1)
    hist = []
    k = 0
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
        held = evaluate(net, ts, n_train, ts.n, batch)
        hist.append((float(np.mean(run)) if run else float("nan"), held))
        log(f"epoch {ep + 1}/{epochs}: train loss {hist[-1][0]:.4f}, held-out loss {held:.4f} "
            f"({time.time() - t0:.1f} s, {nb} steps of {batch})")
    return hist


def evaluate(net: ProposalNet, ts: TrainSet, i0: int, i1: int, batch=256) -> float:
    """Mean loss of samples i0..i1 (forward only)."""
    y = DeviceArray(net.ctx, (batch, 12), np.float32)
    tot, cnt = 0.0, 0
    for o in range(i0, i1, batch):
        b = min(batch, i1 - o)
        net.forward(ts.crops.row(o), b, None, y.ptr)
        e = y.download()[:b].astype(np.float64) - ts.labels_host[o:o + b] / SCALE
        tot += float((e * e).sum())
        cnt += b
    return tot / max(1, cnt)
