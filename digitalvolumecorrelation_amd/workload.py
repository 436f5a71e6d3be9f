"""Algorithmic work of a search, for roofline figures.

The search reads, per point, the voxels that the tricubic interpolant touches: the
4x4x4 voxel support [i-1, i+2]^3 of every cell that holds a sample, once for the
reference cloud and once for the target cloud at its converged warp (the kern_4d
boxes of Interpolate::kernels, Core/Interpolate.cpp:144-258, hold a superset).  That
support, times the bytes per voxel, is the compulsory volume traffic of one point.

The warp restates FloatingCloud::affine_to (Core/FloatingCloud.cpp:236-286) with
the rotation of SearchParams::set_param_vect (Core/SearchParams.cpp:53-107).
"""
from __future__ import annotations

import numpy as np


def sphere_offsets(seed: int, rad: float, num: int, aspect=(1.0, 1.0, 1.0)) -> np.ndarray:
    """Sample offsets of a sphere cloud (FloatingCloud.cpp:51-78): std::mt19937(seed)
    through uniform_real_distribution<double>(-rad, rad), three draws per try, kept when
    inside the ball, then scaled by aspect.  numpy's legacy MT19937 seeds like
    std::mt19937, and libstdc++'s generate_canonical<double, 53> combines two 32-bit
    draws as (d0 + d1 * 2^32) / 2^64."""
    rs = np.random.RandomState(seed)
    out = np.zeros((0, 3))
    while len(out) < num:
        want = 3 * (2 * (num - len(out)) + 22)
        raw = rs.randint(0, 2 ** 32, size=2 * want, dtype=np.uint64).astype(np.float64)
        u = (raw[0::2] + raw[1::2] * 2.0 ** 32) / 2.0 ** 64
        u = np.where(u >= 1.0, np.nextafter(1.0, 0.0), u)
        d = (u * (2 * rad) - rad).reshape(-1, 3)
        out = np.vstack([out, d[np.sqrt((d ** 2).sum(1)) <= rad]])
    return out[:num] * np.asarray(aspect, dtype=np.float64)


def rotation(phi: float, the: float, psi: float) -> np.ndarray:
    cf, sf, ct, st, cp, sp = np.cos(phi), np.sin(phi), np.cos(the), np.sin(the), np.cos(psi), np.sin(psi)
    return np.array([[ct * cf, ct * sf, -st],
                     [sp * st * cf - cp * sf, sp * st * sf + cp * cf, ct * sp],
                     [cp * st * cf + sp * sf, cp * st * sf - sp * cf, ct * cp]])


def warp_cloud(stable: np.ndarray, refpt, par, ndof: int) -> np.ndarray:
    """Moving cloud of a stable cloud [n,3] about refpt under par[ndof]."""
    par = np.asarray(par, dtype=np.float64)
    u = par[:3]
    x = stable + u
    if ndof > 3:
        m = np.asarray(refpt, dtype=np.float64) + u
        r = x - m
        x = x + (r @ rotation(*par[3:6]).T - r)
        if ndof > 6:
            e = par[6:12]
            S = np.array([[e[0], e[3], e[5]], [e[3], e[1], e[4]], [e[5], e[4], e[2]]])
            x = x + (x - m) @ S.T
    return x


def support_voxels(xyz: np.ndarray) -> int:
    """Number of distinct voxels in the [i-1, i+2]^3 supports of the cells of xyz."""
    cell = np.floor(xyz).astype(np.int64)
    lo = cell.min(0) - 1
    hi = cell.max(0) + 2
    ext = hi - lo + 1
    m = np.zeros(ext, dtype=bool)
    c = cell - lo
    m[c[:, 0], c[:, 1], c[:, 2]] = True
    for ax in range(3):
        d = m.copy()
        for s in (-1, 1, 2):
            # voxel v is touched by cell v - s for s in {-1, 0, 1, 2}
            src = [slice(None)] * 3
            dst = [slice(None)] * 3
            if s > 0:
                src[ax] = slice(0, ext[ax] - s)
                dst[ax] = slice(s, None)
            else:
                src[ax] = slice(-s, None)
                dst[ax] = slice(0, ext[ax] + s)
            d[tuple(dst)] |= m[tuple(src)]
        m = d
    return int(m.sum())


def search_bytes(points: np.ndarray, offsets: np.ndarray, results: np.ndarray, ndof: int, bytes_per: int,
                 sample: int | None = None) -> tuple[float, int]:
    """Mean compulsory volume bytes per searched point and the number of points it is
    taken over.  offsets: the sample offsets of one point's cloud about its centre
    (sphere offsets; for a cube, the grid offsets).  Unsearched and failed points are
    skipped (their reads stop early)."""
    good = np.flatnonzero(results["status"] == 0)
    if sample is not None and len(good) > sample:
        good = good[np.linspace(0, len(good) - 1, sample).astype(np.int64)]
    tot = 0.0
    for i in good:
        stable = points[i] + offsets
        ref = support_voxels(stable)
        tar = support_voxels(warp_cloud(stable, points[i], results["par"][i][:ndof], ndof))
        tot += bytes_per * (ref + tar)
    return (tot / len(good) if len(good) else 0.0), len(good)


def cells(xyz: np.ndarray) -> int:
    """Number of distinct unit cells (truncated coordinates, Point.cpp:57-66) holding samples."""
    return int(len(np.unique(np.floor(xyz).astype(np.int64), axis=0)))


def search_flops(points: np.ndarray, offsets: np.ndarray, results: np.ndarray, ndof: int,
                 sample: int | None = None) -> tuple[float, float, int]:
    """Mean algorithmic floating-point work per searched point in the reference's own
    formulation (Interpolate.cpp:377-402, 472-567; Search.cpp:194-246, 490-525): the
    64x64 coefficient product of every distinct cell the reference and target clouds
    touch (2*4096 per cell) and the 64-term evaluation of every sample in every objective
    evaluation (2*64 per sample).  Evaluations per point: one reference sampling, (1 +
    ndof) per Gauss-Newton iteration (the forward-difference Jacobian of LM_prep_at) and
    the final objective (Search.cpp:187).  Target cells are those of the converged cloud
    (a lower bound of the union over the search).  Returns (mean FLOP per point, mean
    evaluations per point, points sampled)."""
    good = np.flatnonzero(results["status"] == 0)
    if sample is not None and len(good) > sample:
        good = good[np.linspace(0, len(good) - 1, sample).astype(np.int64)]
    n = len(offsets)
    tot = ev = 0.0
    for i in good:
        stable = points[i] + offsets
        c = cells(stable) + cells(warp_cloud(stable, points[i], results["par"][i][:ndof], ndof))
        e = 1 + int(results["iters"][i]) * (1 + ndof) + 1
        tot += 2.0 * 4096 * c + 2.0 * 64 * n * e
        ev += e
    m = max(len(good), 1)
    return tot / m, ev / m, len(good)
