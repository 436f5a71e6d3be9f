"""Train the start-warp proposal network on the GPU with the library's own kernels.

    python -m digitalvolumecorrelation_amd.train_proposal [--volumes 17] [--epochs 6] [--out FILE]

Training set: seeded synthetic uint8 pairs of 128^3 (smoothed noise, sigma 1.5-2.5)
under random smooth deformations (translation up to 8 voxels per axis, rotation up to
3 degrees about a random axis through the centre, symmetric strain up to 2 %, plus a
1-voxel sinusoidal field of wavelength 64), a 23^3 grid of points per pair at spacing
3.  Labels are the converged 12-DOF warps of the plain search (serial chain, cube 31
with 2197 samples, ZNSSD, tricubic).  Inputs are the 32^3 reference/target crops
around each point, kept in HBM as BF16.  Writes the FP32 parameter vector and the
per-epoch losses to an .npz file and prints one JSON summary line.
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
    ap.add_argument("--volumes", type=int, default=17)
    ap.add_argument("--size", type=int, default=128)
    ap.add_argument("--spacing", type=float, default=3.0)
    ap.add_argument("--epochs", type=int, default=6)
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
    ts = P.build_train_set(ctx, net, range(a.volume_seed0, a.volume_seed0 + a.volumes), size=a.size,
                           spacing=a.spacing, log=log)
    t1 = time.time()
    log(f"training set: {ts.n} crop pairs in HBM ({ts.n * P.CROP_ELEMS * 2 / 2**30:.1f} GiB), {t1 - t0:.1f} s")
    hist = P.train(net, ts, epochs=a.epochs, batch=a.batch, lr=a.lr, log=log)
    t2 = time.time()
    lab = ts.labels_host
    net.save(a.out, history=np.array(hist), samples=ts.n, volumes=a.volumes, epochs=a.epochs, batch=a.batch,
             label_std=lab.std(axis=0))
    print(json.dumps({"samples": ts.n, "epochs": a.epochs, "batch": a.batch, "data_s": round(t1 - t0, 1),
                      "train_s": round(t2 - t1, 1), "loss": hist, "label_std": lab.std(axis=0).round(4).tolist(),
                      "out": a.out}), flush=True)


if __name__ == "__main__":
    main()
