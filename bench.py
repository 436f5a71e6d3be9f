"""Throughput of the DVC correlation path on one MI355X.

Workload: 256^3 uint16 reference/target pair (smoothed noise, sigma 1.5; target under
an affine map: 0.5 % strain along z, 1 degree about x, translation (2.4, 0.3, -1.1)),
a 15^3 grid of 3375 points at spacing 12, sphere subvolumes of diameter 41 with 5000
samples, ZNSSD, tricubic, 12-DOF Gauss-Newton search.  One step = one full search of
every point (dvc_search_execute: the point loop of Core/dvc.cpp:233-303), volumes and
point tables already resident in HBM.

One step also fits the strain field of the result (the 26 grid neighbours and the
point, strain.cpp:402-591) on the device.

Prints one JSON line:  metric/value (points/s), the roofline of the step (the
coefficient contractions of the reference's formulation against FP32 matrix peak, the
bound of the path; algorithmic bytes against HBM beside it; device time from HIP events on the library's
stream), the learned-proposal
leg with its agreement counts against the plain chain, and the oracle's CPU throughput
on the same processing order (same box, stated thread count) with the agreement of the
GPU run against it point by point.

    python bench.py [--steps K] [--warmup W] [--gpus 1] [--cpu-points M] [--no-cpu]
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

METRIC = "points/s at <=1e-3 voxel agreement: plain GPU vs NN-proposal GPU vs CPU OpenMP"
HBM_PEAK_GBS = 8000.0
FP32_MATRIX_PEAK_TFLOPS = 157.3


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--threads", type=int, default=0, help="threads per point (0 = library default)")
    ap.add_argument("--cpu-points", type=int, default=0,
                    help="oracle sample: first M points of the order (0: every point of the default workload)")
    ap.add_argument("--cpu-threads", type=int, default=0, help="oracle threads (0 = min(16, cores))")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-sweep", type=int, default=0, metavar="M",
                    help="only time the oracle on the first M points of the default workload at 1, 2, 4, 8 and "
                         "16 threads (the reference's thread sweep) and print one JSON line")
    ap.add_argument("--bytes-sample", type=int, default=400, help="points sampled for the algorithmic byte count")
    ap.add_argument("--cache", default="", help="directory to keep the generated volumes in between runs")
    ap.add_argument("--no-proposal", action="store_true", help="skip the learned-proposal leg")
    ap.add_argument("--workload", default="affine", choices=["affine", "full", "learned", "uneven"],
                    help="affine: the 12-DOF 256^3 workload (default); full: the 1536^3 / 10^6-point case; "
                         "learned: the 1024^3 / 10^5-point evaluation of the learned proposal; uneven: 512^3 with "
                         "5*10^4 points of per-point subvolume shape, size and sample count (volumes of these "
                         "made on the device, each reported in its own JSON line)")
    ap.add_argument("--full-size", type=int, default=1536)
    ap.add_argument("--full-grid", type=int, default=100)
    ap.add_argument("--weights", default="", help="proposal network parameters (.npz); default: the package's")
    ap.add_argument("--copy-bytes", type=int, default=0,
                    help="only run a plain device-to-device copy of this many bytes (3 times), the known-size "
                         "kernel that scales the HBM byte counters of a profiler run")
    return ap.parse_args()


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def load_case(cache: str, synth):
    """The synthetic workload; its two volumes optionally kept as .npy under `cache`."""
    if not cache:
        return synth.affine_u16()
    d = Path(cache)
    r, t = d / "affine_u16_256_ref.npy", d / "affine_u16_256_tar.npy"
    if r.exists() and t.exists():
        case = synth.affine_u16(volumes=(np.load(r), np.load(t)))
    else:
        case = synth.affine_u16()
        d.mkdir(parents=True, exist_ok=True)
        np.save(r, case.ref)
        np.save(t, case.tar)
    return case


def agreement(res, plain):
    """Agreement of a result table with the plain chain's, point by point."""
    both = (res["status"] == 0) & (plain["status"] == 0)
    du = np.abs(res["par"][:, :3] - plain["par"][:, :3]).max(axis=1)
    rel = np.zeros(len(res))
    rel[both] = np.abs(res["obj"][both] - plain["obj"][both]) / np.abs(plain["obj"][both])
    return {"status_equal_plain": bool(np.array_equal(res["status"], plain["status"])),
            "points_status_differs": int((res["status"] != plain["status"]).sum()),
            "max_du_vs_plain_voxel": float(du[both].max()) if both.any() else None,
            "points_outside_1e-3_voxel_of_plain": int((both & (du > 1e-3)).sum()),
            "max_obj_rel_vs_plain": float(rel[both].max()) if both.any() else None,
            "points_outside_1e-5_obj_rel_of_plain": int((both & (rel > 1e-5)).sum())}


def proposal_leg(ctx, a, case, params, order, neigh, plain, D):
    """The same workload with learned start warps: every step proposes a warp for each
    point (crops + network forward), searches all points from it in one launch, runs the
    judge pass from the reference's own start guess on those results, takes the judge's
    result where both passes converged and agree, and searches every other point again
    in the serial chain's order (DVC_START_JUDGED).  Its rate is reported with its
    agreement counts against the plain chain, not as a rate at 1e-3 agreement unless
    every point is within it."""
    from digitalvolumecorrelation_amd import proposal as P
    wpath = Path(a.weights) if a.weights else P.WEIGHTS
    if not wpath.exists():
        log(f"no proposal weights at {wpath}: NN-proposal leg skipped")
        return None
    net = P.ProposalNet(ctx, max_batch=4096, params=P.ProposalNet.load_params(wpath))
    ctx.set_start_mode(D.START_JUDGED)
    ctx.prepare(params, case.points, order, neigh)
    ctx.set_start_mode(D.START_CHAIN)
    net.propose()
    tuned = ctx.autotune(2) if not a.threads else (ctx.block_threads(), None)
    log(f"NN proposal: threads per point {tuned[0]}")
    for _ in range(a.warmup):
        net.propose()
        ctx.execute()
        ctx.kernel_ms()
    t0 = time.perf_counter()
    for _ in range(a.steps):
        net.propose()
        ctx.execute()
    ctx.kernel_ms()                       # the last step's closing event
    wall_ms = (time.perf_counter() - t0) * 1e3 / a.steps
    res = ctx.fetch()
    prop, judge, kept = ctx.fetch_passes()
    n = len(case.points)
    log(f"NN proposal: {wall_ms:.2f} ms/step")
    pg = prop["status"] == 0
    agr = agreement(res, plain)
    within = agr["points_outside_1e-3_voxel_of_plain"] == 0 and agr["points_status_differs"] == 0
    return {"value": n / (wall_ms / 1e3), "unit": "points/s", "ms_per_step": wall_ms,
            "speedup_vs_plain": None, "weights": wpath.name, "threads_per_point": tuned[0],
            "mode": "judged: proposal pass, judge pass from the reference start guess over its results, the judge's "
                    "result where both converge and agree (1e-3 voxel, 1e-5 obj), repair of every other point in "
                    "the serial chain's order",
            "judge_results": int((kept == 0).sum()), "repaired": int((kept == 2).sum()),
            "rate_at_1e-3_agreement": bool(within),
            "mean_iters_proposal_pass": float(prop["iters"][pg].mean()) if pg.any() else None,
            "mean_iters_plain": float(plain["iters"][plain["status"] == 0].mean()),
            **agr}


def host_cpu():
    """The host's processor as lscpu reports it, and the cores this process may use."""
    info = {}
    try:
        import subprocess
        out = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=10).stdout
        for line in out.splitlines():
            k, _, v = line.partition(":")
            if k.strip() in ("Model name", "Socket(s)", "Core(s) per socket", "Thread(s) per core", "CPU(s)"):
                info[k.strip()] = v.strip()
    except Exception as e:                      # noqa: BLE001 - lscpu absent: say so
        info["lscpu"] = f"unavailable ({e})"
    try:
        info["cpus_available_to_process"] = len(os.sched_getaffinity(0))
    except AttributeError:
        info["cpus_available_to_process"] = os.cpu_count()
    return info


def uneven(a, D, synth):
    """The batched uneven workload: every point its own subvolume shape, size and sample
    count (dvc_set_point_geometry), SSD and ZNSSD, 6-DOF.  Per objective: the serial chain
    with fixed and autotuned threads per point, the judged proposal mode and the one-launch
    search from one rigid start; agreement with the chain; the oracle on a prefix of the
    order (one RunControl per point, volumes copied back to the host)."""
    from digitalvolumecorrelation_amd import proposal as P
    t0 = time.time()
    ctx = D.Context(0)
    case, geom, sub, nsamp = synth.uneven_u8(ctx)
    n = len(case.points)
    log(f"volumes made on the device in {time.time() - t0:.1f} s; {n} points, samples per point "
        f"{nsamp.min()}..{nsamp.max()} (cubes rounded up), {len(set(zip(geom, sub, nsamp)))} geometries")
    order, neigh = D.knn_order_neighbors(case.points, start=case.start)
    net = P.ProposalNet(ctx, max_batch=4096, params=P.ProposalNet.load_params(Path(a.weights) if a.weights else P.WEIGHTS))
    out = {"workload": f"batched uneven: {case.shape[0]}^3 u8 pair made on the device (sinusoids of amplitude 3, "
                       f"wavelength 128), {n} random points, each a cube or sphere of size 21..61 with 1000..10000 "
                       f"samples, tricubic, 6-DOF, disp_max 6", "points": n}
    ref = tar = None
    for obj in ("znssd", "ssd"):
        params = D.make_params(**dict(case.params, obj=obj))

        def timed(mode, starts=None, propose=None, threads=512, tune=False):
            ctx.set_point_geometry(geom, sub, nsamp)
            ctx.set_start_mode(mode)
            ctx.set_block_threads(threads)
            ctx.set_slice_launches(0, 0, 0)       # the default shape; the autotuner may change it
            t1 = time.time()
            ctx.prepare(params, case.points, order, neigh)
            prep = time.time() - t1
            if starts is not None:
                ctx.set_starts(starts)
            if propose:
                propose()
            tuned = ctx.autotune(1) if tune else None
            for _ in range(max(1, a.warmup)):
                if propose:
                    propose()
                ctx.execute()
                ctx.kernel_ms()
            t1 = time.perf_counter()
            for _ in range(a.steps):
                if propose:
                    propose()
                ctx.execute()
            ctx.kernel_ms()
            ms = (time.perf_counter() - t1) * 1e3 / a.steps
            return ms, ctx.fetch(), prep, tuned

        ms_p, plain, prep, _ = timed(D.START_CHAIN)
        levels = ctx.levels()
        fixed = {}
        for t in (128, 256, 1024):
            fixed[t] = timed(D.START_CHAIN, threads=t)[0]
        fixed[512] = ms_p
        ms_t, tuned_r, _, tuned = timed(D.START_CHAIN, tune=True)
        ms_j, judged, _, _ = timed(D.START_JUDGED, propose=net.propose)
        kept = ctx.fetch_passes()[2]
        rigid = np.zeros((n, 12))
        rigid[:, :3] = case.params["rigid_trans"]
        ms_r, flat, _, _ = timed(D.START_GIVEN, starts=rigid)
        ok = plain["status"] == 0
        err = np.abs(plain["par"][ok, :3] - case.truth[ok]).max(axis=1)

        def agree(r):
            good = ok & (r["status"] == 0)
            du = np.abs(r["par"][:, :3] - plain["par"][:, :3]).max(axis=1)
            dobj = r["obj"][good] - plain["obj"][good]
            e = np.abs(r["par"][good, :3] - case.truth[good]).max(axis=1)
            return {"good": int((r["status"] == 0).sum()), "within_1e-3_of_plain": int((good & (du <= 1e-3)).sum()),
                    "obj_not_worse_than_plain_by_1e-6": int((dobj <= 1e-6).sum()),
                    "obj_excess_vs_plain_p50_p99_max": [float(np.quantile(dobj, 0.5)), float(np.quantile(dobj, 0.99)),
                                                        float(dobj.max())],
                    "median_err_vs_imposed_voxel": float(np.median(e)),
                    "mean_evaluations": float(r["iters"][r["status"] == 0].mean())}

        leg = {"levels": levels, "prepare_s": round(prep, 1),
               "plain": {"value": n / (ms_p / 1e3), "ms_per_step": ms_p, "threads_per_point": 512, "good": int(ok.sum()),
                         "mean_evaluations": float(plain["iters"][ok].mean()),
                         "median_err_vs_imposed_voxel": float(np.median(err)),
                         "max_err_vs_imposed_voxel": float(err.max())},
               "plain_fixed_threads_ms_per_step": {str(t): fixed[t] for t in sorted(fixed)},
               "plain_autotuned": {"value": n / (ms_t / 1e3), "ms_per_step": ms_t, "threads_per_point": tuned[0],
                                   **agree(tuned_r)},
               "nn_judged": {"value": n / (ms_j / 1e3), "ms_per_step": ms_j, "speedup_vs_plain": ms_p / ms_j,
                             "repaired": int((kept == 2).sum()), **agree(judged)},
               "rigid_start_given": {"value": n / (ms_r / 1e3), "ms_per_step": ms_r, **agree(flat)}}
        log(f"{obj}: plain {ms_p:.1f} ms ({levels} levels), autotuned {ms_t:.1f} ms ({tuned[0]} threads), "
            f"judged {ms_j:.1f} ms, rigid start {ms_r:.1f} ms")
        if not a.no_cpu:
            sys.path.insert(0, str(ROOT / "oracle"))
            import oracle as O
            if ref is None:
                ref = ctx.download(0, case.shape, np.uint8)
                tar = ctx.download(1, case.shape, np.uint8)
            M = min(a.cpu_points or 384, n)
            thr = a.cpu_threads or min(16, os.cpu_count() or 1)
            run = O.make_run(shape=case.shape, bytes_per=1, jacobian=0, **dict(case.params, obj=obj))
            ores, sec, m = O.dvc_run_geo(run, ref, tar, case.points, order, neigh, geom, sub, nsamp, nproc=M,
                                         nthreads=thr)
            idx = order[:m]
            g, o = plain[idx], ores[idx]
            okk = o["status"] == 0
            leg["cpu_baseline"] = {"value": m / sec, "unit": "points/s", "cores": thr, "kind": "port",
                                   "sample": f"first {m} points of the processing order, {sec:.1f} s",
                                   "status_equal": bool(np.array_equal(g["status"], o["status"])),
                                   "iters_equal": int((g["iters"] == o["iters"]).sum()),
                                   "max_du_voxel": float(np.abs(g["par"][okk, :3] - o["par"][okk, :3]).max())}
        out[obj] = leg
    ctx.set_point_geometry()
    ctx.close()
    print(json.dumps(out), flush=True)


def cpu_sweep(a, case, order, neigh):
    """The oracle's points/s on a prefix of the default workload's processing order at
    1..P threads (OpenMP inside interpolation, as the reference parallelises; its thread
    sweep is tests.cpp:215-314)."""
    sys.path.insert(0, str(ROOT / "oracle"))
    import oracle as O
    cores = os.cpu_count() or 1
    M = min(a.cpu_sweep, len(case.points))
    run = O.make_run(shape=case.ref.shape, bytes_per=case.ref.dtype.itemsize, jacobian=0, **case.params)
    rows = []
    for thr in (1, 2, 4, 8, 16):
        if thr > cores:
            break
        _, sec, n = O.dvc_run(run, case.ref, case.tar, case.points, order, neigh, nproc=M, nthreads=thr)
        rows.append({"threads": thr, "points_per_s": n / sec, "seconds": round(sec, 2)})
        log(f"oracle, {thr} threads: {n / sec:.2f} points/s ({sec:.1f} s for {n} points)")
    base = rows[0]["points_per_s"]
    for r in rows:
        r["speedup"] = round(r["points_per_s"] / base, 2)
    print(json.dumps({"cpu_sweep": rows, "sample": f"first {M} points of the processing order", "cores": cores,
                      "workload": "plain 12-DOF affine, 256^3 u16, sphere d=41 / 5000 samples"}), flush=True)


def learned_eval(a, D, synth):
    """The learned proposal's evaluation on a 1024^3 uint8 pair (translation up to 10
    voxels per axis plus a sinusoidal field, 10^5 random points): the plain serial
    chain, the judged proposal mode, the proposal pass alone, and the same one-launch
    search from one rigid start for every point (the method without the network):
    points/s, Gauss-Newton evaluations, fallback rate and agreement with the plain
    search."""
    from digitalvolumecorrelation_amd import proposal as P
    t0 = time.time()
    ctx = D.Context(0)
    case = synth.learned_eval_u8(ctx)
    log(f"volumes made on the device in {time.time() - t0:.1f} s: {case.shape} uint8")
    order, neigh = D.knn_order_neighbors(case.points, start=case.start)
    params = D.make_params(**case.params)
    n = len(case.points)

    def timed(prepare_mode, starts=None, propose=None):
        ctx.set_start_mode(prepare_mode)
        ctx.prepare(params, case.points, order, neigh)
        ctx.set_start_mode(D.START_CHAIN)
        if a.threads:
            ctx.set_block_threads(a.threads)
        if starts is not None:
            ctx.set_starts(starts)
        for _ in range(max(1, a.warmup)):
            if propose:
                propose()
            ctx.execute()
            ctx.kernel_ms()
        t = time.perf_counter()
        for _ in range(a.steps):
            if propose:
                propose()
            ctx.execute()
        ctx.kernel_ms()
        ms = (time.perf_counter() - t) * 1e3 / a.steps
        return ms, ctx.fetch()

    ms_plain, plain = timed(D.START_CHAIN)
    levels = ctx.levels()
    log(f"plain chain: {ms_plain:.1f} ms per step, {levels} levels")
    net = P.ProposalNet(ctx, max_batch=4096, params=P.ProposalNet.load_params(Path(a.weights) if a.weights else P.WEIGHTS))
    ms_j, judged = timed(D.START_JUDGED, propose=net.propose)
    prop, judge, kept = ctx.fetch_passes()
    log(f"judged proposals: {ms_j:.1f} ms per step")
    ms_g, given = timed(D.START_GIVEN, propose=net.propose)
    rigid = np.zeros((n, 12))
    rigid[:, :3] = case.params["rigid_trans"]
    ms_r, flat = timed(D.START_GIVEN, starts=rigid)
    ok = plain["status"] == 0

    def agree(r):
        good = ok & (r["status"] == 0)
        du = np.abs(r["par"][:, :3] - plain["par"][:, :3]).max(axis=1)
        return {"good": int((r["status"] == 0).sum()), "within_1e-3_of_plain": int((good & (du <= 1e-3)).sum()),
                "within_1e-2_of_plain": int((good & (du <= 1e-2)).sum()),
                "mean_evaluations": float(r["iters"][r["status"] == 0].mean()) if (r["status"] == 0).any() else None}

    err = np.abs(plain["par"][ok, :3] - case.truth[ok]).max(axis=1)
    out = {"workload": f"learned-proposal evaluation: {case.shape[0]}^3 u8 pair made on the device (translation "
                       f"{np.round(case.bumps[0, 4:7], 2).tolist()} + sinusoids of amplitude 2, wavelength 256), "
                       f"{n} random points, cube 31 / 2197 samples, ZNSSD, 12-DOF, disp_max 12",
           "levels": levels,
           "plain": {"value": n / (ms_plain / 1e3), "ms_per_step": ms_plain, "good": int(ok.sum()),
                     "mean_evaluations": float(plain["iters"][ok].mean()),
                     "median_err_vs_imposed_voxel": float(np.median(err)), "max_err_vs_imposed_voxel": float(err.max())},
           "nn_judged": {"value": n / (ms_j / 1e3), "ms_per_step": ms_j, "speedup_vs_plain": ms_plain / ms_j,
                         "judge_results": int((kept == 0).sum()), "repaired": int((kept == 2).sum()),
                         "proposal_pass": agree(prop), "selected": agree(judged)},
           "nn_given_only": {"value": n / (ms_g / 1e3), "ms_per_step": ms_g, **agree(given)},
           "rigid_start_given": {"value": n / (ms_r / 1e3), "ms_per_step": ms_r, **agree(flat)}}
    ctx.close()
    print(json.dumps(out), flush=True)


def full_scale(a, D, synth, workload):
    """The full-scale case on one GPU: both volumes made on the device, every point of
    the grid searched by the serial chain (plain method) and by the judged proposal mode,
    the strain field of the plain result, accuracy against the imposed field, and the
    oracle on a prefix of the processing order (volumes copied back to the host)."""
    t0 = time.time()
    ctx = D.Context(0)
    size, n = a.full_size, a.full_grid
    case = synth.full_scale_u16(ctx, size=size, n=n, spacing=14.0 * 100 / n if size == 1536 else (size - 150) / (n - 1))
    t_gen = time.time() - t0
    log(f"volumes made on the device in {t_gen:.1f} s: {case.shape} uint16")
    t1 = time.time()
    order, neigh = D.knn_order_neighbors(case.points, start=case.start)
    t_knn = time.time() - t1
    log(f"order and 75 neighbours of {len(case.points)} points in {t_knn:.1f} s")
    params = D.make_params(**case.params)
    if a.threads:
        ctx.set_block_threads(a.threads)
    t1 = time.time()
    ctx.prepare(params, case.points, order, neigh)
    t_prep = time.time() - t1
    levels = ctx.levels()
    log(f"prepared in {t_prep:.1f} s: {levels} chain launches")
    ctx.execute()
    ms_first = ctx.kernel_ms()
    dev = []
    t1 = time.perf_counter()
    for _ in range(a.steps):
        ctx.execute()
        dev.append(ctx.kernel_ms())
    wall = (time.perf_counter() - t1) * 1e3 / a.steps
    res = ctx.fetch()
    npts = len(case.points)
    good = res["status"] == 0
    err = np.abs(res["par"][good, :3] - case.truth[good]).max(axis=1)
    log(f"plain: {wall:.1f} ms per step ({npts / wall * 1e3:.0f} points/s), good {good.sum()}")
    ctx.strain(case.points, res["par"][:, :3].copy(), res["status"], res["obj"], neigh, ndp=27)
    strain_ms = ctx.kernel_ms()
    # cube cloud offsets about the point (FloatingCloud.cpp:84-128): g^3 grid over [-rad, rad]
    g = int(np.ceil(case.params["subvol_npts"] ** (1.0 / 3.0)))
    rad = case.params["subvol_size"] / 2.0
    k = -rad + np.arange(g) * (2.0 * rad / (g - 1))
    kz, ky, kx = np.meshgrid(k, k, k, indexing="ij")
    offs = np.stack([kx.ravel(), ky.ravel(), kz.ravel()], axis=1)
    per_pt, ns = workload.search_bytes(case.points, offs, res, 12, 2, sample=a.bytes_sample)
    roof = None
    if per_pt:
        achieved = per_pt * int(good.sum()) / (np.mean(dev) / 1e3) / 1e9
        roof = {"bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": achieved / HBM_PEAK_GBS, "algorithmic_bytes_per_point": round(per_pt, 1)}
    nn = None
    if not a.no_proposal:
        nn = proposal_leg(ctx, a, case, params, order, neigh, res, D)
    cpu = None
    if not a.no_cpu:
        sys.path.insert(0, str(ROOT / "oracle"))
        import oracle as O
        ref = ctx.download(0, case.shape, np.uint16)
        tar = ctx.download(1, case.shape, np.uint16)
        thr = a.cpu_threads or min(16, os.cpu_count() or 1)
        M = min(max(a.cpu_points, 2000), npts)       # a prefix of at least 2000 points at this size
        run = O.make_run(shape=case.shape, bytes_per=2, jacobian=0, **case.params)
        ores, sec, m = O.dvc_run(run, ref, tar, case.points, order, neigh, nproc=M, nthreads=thr)
        idx = order[:m]
        ok = ores["status"][idx] == 0
        du = np.abs(res["par"][idx][ok, :3] - ores["par"][idx][ok, :3]).max() if ok.any() else 0.0
        rel = (np.abs(res["obj"][idx][ok] - ores["obj"][idx][ok]) / ores["obj"][idx][ok]).max() if ok.any() else 0.0
        cpu = {"value": m / sec, "unit": "points/s", "cores": thr, "kind": "port",
               "sample": f"first {m} points of the processing order, {sec:.1f} s",
               "agreement": {"status_equal": bool(np.array_equal(res["status"][idx], ores["status"][idx])),
                             "iters_equal": int((res["iters"][idx] == ores["iters"][idx]).sum()),
                             "max_du_voxel": float(du), "max_obj_rel": float(rel)}}
        del ref, tar
    out = {"workload": f"full scale: {size}^3 u16 pair made on the device (8 Gaussian bumps of 4 voxels, width 200, "
                       f"+1% volumetric strain), {n}^3 grid ({npts} points), cube 41 / 8000 samples, ZNSSD, "
                       f"tricubic, 12-DOF, disp_max 10",
           "points": npts, "good": int(good.sum()), "levels": levels,
           "plain": {"value": npts / (wall / 1e3), "unit": "points/s", "ms_per_step": wall,
                     "device_ms_per_step": float(np.mean(dev)), "first_execute_ms": ms_first},
           "nn_proposal": nn, "strain_fit_device_ms": strain_ms,
           "max_err_vs_imposed_voxel": float(err.max()) if err.size else None,
           "median_err_vs_imposed_voxel": float(np.median(err)) if err.size else None,
           "setup_s": {"volumes": round(t_gen, 1), "knn": round(t_knn, 1), "prepare": round(t_prep, 1)},
           "roofline": roof, "cpu_baseline": cpu}
    ctx.close()
    print(json.dumps(out), flush=True)


def main():
    a = parse()
    if a.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print(f"bench.py measures one GPU only: the search runs on a single device (asked for {a.gpus}); "
              "run it without --gpus or with --gpus 1", file=sys.stderr)
        sys.exit(2)

    if a.copy_bytes:
        import torch
        src = torch.ones(a.copy_bytes, dtype=torch.uint8, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes": a.copy_bytes, "launches": 3}), flush=True)
        return

    import digitalvolumecorrelation_amd as D
    from digitalvolumecorrelation_amd import synth, workload

    if a.workload == "full":
        return full_scale(a, D, synth, workload)
    if a.workload == "learned":
        return learned_eval(a, D, synth)
    if a.workload == "uneven":
        return uneven(a, D, synth)

    t0 = time.time()
    case = load_case(a.cache, synth)
    p = case.params
    order, neigh = D.sort_order_neighbors(case.points, case.start)
    if a.cpu_sweep:
        return cpu_sweep(a, case, order, neigh)
    log(f"workload built in {time.time() - t0:.1f} s: {len(case.points)} points, volumes {case.ref.shape} "
        f"{case.ref.dtype}")

    ctx = D.Context(0)
    ctx.upload(0, case.ref)
    ctx.upload(1, case.tar)
    if a.threads:
        ctx.set_block_threads(a.threads)
    params = D.make_params(**p)
    ctx.set_search_strain(ndp=27)           # the strain field of every step's result, on the device
    ctx.prepare(params, case.points, order, neigh)
    tuned = None
    if not a.threads:
        tuned = ctx.autotune(2)
        log(f"autotuned: {tuned[0]} threads per point, slice launches {ctx.slice_info()} ({tuned[1]:.2f} ms per step)")

    for _ in range(a.warmup):
        ctx.execute()
        ctx.kernel_ms()
    log(f"warmup done ({a.warmup} steps)")

    dev_ms = []
    t_start = time.perf_counter()
    for _ in range(a.steps):
        ctx.execute()
        dev_ms.append(ctx.kernel_ms())       # waits for the step's closing event
    t_end = time.perf_counter()
    wall_ms = (t_end - t_start) * 1e3 / a.steps
    res = ctx.fetch()
    fit = ctx.fetch_search_strain()
    npts = len(case.points)
    value = npts / (wall_ms / 1e3)
    sliced = ctx.slice_info()
    chain_launches = ctx.levels()
    threads_pp = ctx.block_threads()
    log(f"timed {a.steps} steps: {wall_ms:.2f} ms/step wall, {np.mean(dev_ms):.2f} ms/step device")

    # the device strain fit equals the library's stand-alone fit of the fetched results
    host_fit = ctx.strain(case.points, res["par"][:, :3].copy(), res["status"], res["obj"], neigh, ndp=27)
    strain_equal = all(np.array_equal(fit[f], host_fit[f]) for f in ("disp", "npts", "E", "L", "DG"))

    good = res["status"] == 0
    truth_err = float(np.abs(res["par"][good, :3] - case.truth[good]).max()) if good.any() else None

    # roofline of the step: compulsory volume bytes (HBM bound) and the reference
    # formulation's floating-point work (the 64x64 coefficient products and 64-term
    # evaluations of the tricubic Lekien-Marsden interpolant, against FP32 matrix peak)
    offs = workload.sphere_offsets(params.sphere_seed, p["subvol_size"] / 2.0, p["subvol_npts"])
    per_pt, nsamp = workload.search_bytes(case.points, offs, res, p["ndof"], case.ref.dtype.itemsize,
                                          sample=a.bytes_sample)
    flop_pt, evals_pt, _ = workload.search_flops(case.points, offs, res, p["ndof"], sample=a.bytes_sample)
    step_bytes = per_pt * int(good.sum())
    dev_s = float(np.mean(dev_ms)) / 1e3
    achieved = step_bytes / dev_s / 1e9
    # the coefficient contractions (2*4096 per distinct cell) are the matrix-unit part of
    # the reference formulation; the 64-term evaluations (2*64 per sample and evaluation)
    # are vector work
    contr_pt = flop_pt - 2.0 * 64 * len(offs) * evals_pt
    tflops = contr_pt * int(good.sum()) / dev_s / 1e12
    tflops_all = flop_pt * int(good.sum()) / dev_s / 1e12
    traffic = valu = None
    tf = ROOT / "profiles" / "slice_kernel_affine_u16_256_traffic.json"
    if tf.exists():
        prof = json.loads(tf.read_text())
        traffic = prof.get("hbm_bytes_per_step")
        valu = prof.get("valu_busy_on_active_cus")
    own_evals = 1.0 + float(res["iters"][good].mean()) if good.any() else None
    # the bound of the path in the reference's formulation is FP32 compute (about 400
    # FLOP per compulsory byte, far above the ~20 FLOP/B ridge of the chip): the line's
    # roofline is that one, with the HBM figures beside it
    roof = {"bound": "mfma", "achieved": round(tflops, 3), "peak": FP32_MATRIX_PEAK_TFLOPS, "unit": "TFLOP/s",
            "frac": tflops / FP32_MATRIX_PEAK_TFLOPS, "traffic": traffic,
            "traffic_per": "step (HBM bytes of one step, FETCH_SIZE + WRITE_SIZE scaled by a 1 GiB copy, from "
                           "profiles/slice_kernel_affine_u16_256_traffic.json)",
            "kernels": "slice_kernel<2,512> (narrow dependency levels), resume_kernel, search_kernel<2,B>",
            "launches_per_step": sliced[0] * (sliced[2] + 1) + (chain_launches - sliced[0]),
            "flop_alg_contraction_per_point": round(contr_pt),
            "flop_alg_per_point": round(flop_pt),
            "achieved_all_flop_alg": round(tflops_all, 3),
            "definition": "achieved: the contraction part of FLOP_alg, 2*4096 per distinct cell of the reference "
                          "and converged target clouds (the 64x64 coefficient products), per second; FLOP_alg adds "
                          "2*64 per sample per evaluation, evaluations = 1 + iterations*(1+ndof) + 1 "
                          "(forward-difference Jacobian), as the reference computes them",
            "evaluations_per_point_reference": round(evals_pt, 2),
            "evaluations_per_point_own": round(own_evals, 2) if own_evals else None,
            "valu_busy_on_active_cus": valu,
            "latency_note": "the serial chain's dependency levels (start guesses from earlier neighbours) bound "
                            "the step: each level waits for the one before; HBM and matrix units are far "
                            "from saturated by construction",
            "hbm": {"achieved": round(achieved, 3), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                    "frac": achieved / HBM_PEAK_GBS, "algorithmic_bytes_per_point": round(per_pt, 1),
                    "bytes_points_sampled": nsamp}}

    nn = None
    if not a.no_proposal:
        nn = proposal_leg(ctx, a, case, params, order, neigh, res, D)
        if nn:
            nn["speedup_vs_plain"] = nn["value"] / value

    cpu = None
    agree = None
    if not a.no_cpu:
        sys.path.insert(0, str(ROOT / "oracle"))
        import oracle as O
        hc = host_cpu()
        thr = a.cpu_threads or min(16, hc.get("cpus_available_to_process") or 1)
        M = min(a.cpu_points, npts) if a.cpu_points else npts
        run = O.make_run(shape=case.ref.shape, bytes_per=case.ref.dtype.itemsize, jacobian=0, **p)
        log(f"oracle: first {M} points of the order, {thr} threads")
        ores, sec, n = O.dvc_run(run, case.ref, case.tar, case.points, order, neigh, nproc=M, nthreads=thr)
        cpu = {"value": n / sec, "unit": "points/s", "cores": thr, "kind": "port",
               "sample": f"{'every point' if n == npts else 'the first ' + str(n) + ' points'} of the "
                         f"processing order of the same workload ({n} points, {sec:.1f} s): serial point loop with "
                         f"{thr} OpenMP threads inside interpolation, forward-difference Jacobian (h = 1e-10)",
               "host": hc,
               "notes": "cores: the threads used; a one-GPU job on this host is granted 16 cores. The port reads "
                        "volumes from memory, where the reference reads each point's two boxes from disk "
                        "(Interpolate.cpp:151-193), and parallelises the uint16 conversion it runs serially "
                        "(Interpolate.cpp:219-250): both make this baseline faster than the reference."}
        idx = order[:n]
        g, o = res[idx], ores[idx]
        ok = o["status"] == 0
        du = float(np.abs(g["par"][ok, :3] - o["par"][ok, :3]).max()) if ok.any() else 0.0
        rel = float((np.abs(g["obj"][ok] - o["obj"][ok]) / np.abs(o["obj"][ok])).max()) if ok.any() else 0.0
        agree = {"points": int(n), "status_equal": bool(np.array_equal(g["status"], o["status"])),
                 "iters_equal": int((g["iters"] == o["iters"]).sum()), "max_du_voxel": du, "max_obj_rel": rel,
                 "within_1e-3": bool(du <= 1e-3)}

    out = {
        "metric": METRIC,
        "value": value,
        "unit": "points/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": wall_ms,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": "plain 12-DOF affine, 256^3 u16 pair, 15^3 grid (3375 points) at spacing 12, "
                               "sphere d=41 / 5000 samples, ZNSSD, tricubic, disp_max 10, then the 27-point "
                               "strain field of the result",
                   "points": npts, "good": int(good.sum()), "threads_per_point": threads_pp,
                   "slice_launches": {"chain_launches_sliced": sliced[0], "workgroups_per_point": sliced[1],
                                      "slice_launches_per_chain_launch": sliced[2],
                                      "chain_launches": chain_launches},
                   "device_ms_per_step": float(np.mean(dev_ms)), "max_err_vs_imposed_voxel": truth_err,
                   "strain_fit": "inside the step (device); equal to the stand-alone fit of the fetched results: "
                                 f"{strain_equal}",
                   "arith": "f32 interpolation and sums, f64 warp, normal equations and solve"},
        "roofline": roof,
        "nn_proposal": nn,
        "cpu_baseline": cpu,
        "agreement_vs_cpu": agree,
    }
    ctx.close()
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
