"""Train the start-warp proposal network on the GPU with the library's own kernels.

    python -m digitalvolumecorrelation_amd.train_proposal [--pairs 48] [--epochs 30] [--out FILE]

Training set: seeded synthetic uint8 pairs of 384^3 (smoothed noise, sigma 1.5-2.5).
Each pair's target is the reference warped tile by tile on the device (64^3 tiles,
216 per pair), every tile under its own random affine map about its centre:
translation up to 8 voxels per axis, rotation up to 3 degrees about a random axis,
symmetric strain up to 2 %.  Each reference texture serves several pairs with
different maps, so no texture stands for one deformation.  3^3 points at spacing 5
around every tile centre.  Labels are the converged 12-DOF warps of the plain search
(cube 21 with 1331 samples, ZNSSD, tricubic, each point started from its tile's
imposed translation).  Inputs are the 32^3 reference/target crops around each point,
kept in HBM as BF16.  The held-out loss comes from pairs on a texture never trained
on; the epoch with the lowest held-out loss is kept.  Writes the FP32 parameter
vector and the per-epoch losses to an .npz file and prints one JSON summary line.
"""
from __future__ import annotations

import argparse
import json
import sys
import time

import numpy as np

from . import proposal as P
from ._lib import Context


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--pairs", type=int, default=32)
    ap.add_argument("--maps-per-ref", type=int, default=8, help="pairs sharing one reference texture")
    ap.add_argument("--per-axis", type=int, default=3, help="points per tile and axis")
    ap.add_argument("--spacing", type=float, default=5.0)
    ap.add_argument("--held-pairs", type=int, default=2)
    ap.add_argument("--trans", type=float, default=8.0, help="largest translation per axis of a tile's map")
    ap.add_argument("--device-textures", action="store_true",
                    help="make every reference texture on the device (dvc_synth_noise), one per pair")
    ap.add_argument("--size", type=int, default=384)
    ap.add_argument("--epochs", type=int, default=20)
    ap.add_argument("--batch", type=int, default=256)
    ap.add_argument("--lr", type=float, default=1e-3)
    ap.add_argument("--seed", type=int, default=0, help="initial weights")
    ap.add_argument("--volume-seed0", type=int, default=100)
    ap.add_argument("--out", default=str(P.WEIGHTS))
    a = ap.parse_args(argv)

    def log(m):
        print(m, file=sys.stderr, flush=True)

    t0 = time.time()
    ctx = Context(0)
    net = P.ProposalNet(ctx, max_batch=a.batch, training=True, params=P.init_params(a.seed))
    s0 = a.volume_seed0
    kw = dict(size=a.size, log=log, maps_per_ref=a.maps_per_ref, per_axis=a.per_axis, spacing=a.spacing,
              device_textures=a.device_textures, trans=a.trans)
    ts = P.build_train_set(ctx, net, range(s0, s0 + a.pairs), **kw)
    held = P.build_train_set(ctx, net, range(s0 + 1000, s0 + 1000 + a.held_pairs), **kw)
    t1 = time.time()
    log(f"training set: {ts.n} crop pairs in HBM ({ts.n * P.CROP_ELEMS * 2 / 2**30:.1f} GiB), held out {held.n}, "
        f"{t1 - t0:.1f} s")
    hist = P.train(net, ts, held, epochs=a.epochs, batch=a.batch, lr=a.lr, log=log)
    t2 = time.time()
    per = P.evaluate(net, held, 0, held.n, a.batch, per_output=True)
    rmse = (np.sqrt(per) * P.SCALE).round(5).tolist()
    lab = ts.labels_host
    net.save(a.out, history=np.array(hist), samples=ts.n, pairs=a.pairs, maps_per_ref=a.maps_per_ref,
             epochs=a.epochs, batch=a.batch, trans=a.trans, device_textures=a.device_textures,
             label_std=lab.std(axis=0), held_rmse=np.array(rmse))
    print(json.dumps({"samples": ts.n, "held_out": held.n, "epochs": a.epochs, "batch": a.batch,
                      "data_s": round(t1 - t0, 1), "train_s": round(t2 - t1, 1), "loss": hist,
                      "held_rmse_per_parameter": rmse, "label_std": lab.std(axis=0).round(4).tolist(),
                      "out": a.out}), flush=True)


if __name__ == "__main__":
    main()
