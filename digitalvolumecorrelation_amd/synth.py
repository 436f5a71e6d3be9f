"""Seeded synthetic volume pairs with imposed deformations (ground truth known).

Reference volume: uniform noise, Gaussian-smoothed (sigma voxels, truncate 4 sigma,
reflect), min-max rescaled to the full uint8 / uint16 range.  Target volume: the
reference resampled with Catmull-Rom (the tricubic interpolant of the path) at the
inverse warp, target(x) = ref(phi^-1(x)), rounded.  The displacement of a point X
is u(X) = phi(X) - X.  Volumes are indexed vol[z, y, x]; points are (x, y, z).
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np


def smoothed_noise(shape, sigma: float, dtype=np.uint8, seed: int = 0) -> np.ndarray:
    from scipy.ndimage import gaussian_filter
    rng = np.random.default_rng(seed)
    v = rng.random(shape, dtype=np.float32)
    v = gaussian_filter(v, sigma, mode="reflect", truncate=4.0)
    lo, hi = float(v.min()), float(v.max())
    top = float(np.iinfo(dtype).max)
    return np.rint((v - lo) / (hi - lo) * top).astype(dtype)


def _cr_w(t: np.ndarray) -> np.ndarray:
    t2 = t * t
    t3 = t2 * t
    return np.stack([0.5 * (-t3 + 2 * t2 - t), 0.5 * (3 * t3 - 5 * t2 + 2),
                     0.5 * (-3 * t3 + 4 * t2 + t), 0.5 * (t3 - t2)], axis=-1)


def catmull_rom(vol: np.ndarray, xyz: np.ndarray, chunk: int = 1 << 20) -> np.ndarray:
    """Catmull-Rom value of vol at points xyz (N,3), indices clamped to the volume."""
    T, H, W = vol.shape
    flat = vol.reshape(-1).astype(np.float64)
    out = np.empty(len(xyz))
    for s in range(0, len(xyz), chunk):
        p = xyz[s:s + chunk]
        i = np.floor(p).astype(np.int64)
        t = p - i
        wx, wy, wz = _cr_w(t[:, 0]), _cr_w(t[:, 1]), _cr_w(t[:, 2])
        acc = np.zeros(len(p))
        for kz in range(4):
            z = np.clip(i[:, 2] - 1 + kz, 0, T - 1) * (H * W)
            for ky in range(4):
                y = np.clip(i[:, 1] - 1 + ky, 0, H - 1) * W
                wyz = wz[:, kz] * wy[:, ky]
                row = np.zeros(len(p))
                for kx in range(4):
                    x = np.clip(i[:, 0] - 1 + kx, 0, W - 1)
                    row += wx[:, kx] * flat[z + y + x]
                acc += wyz * row
        out[s:s + chunk] = acc
    return out


def _shift_axis(v: np.ndarray, shift: float, axis: int) -> np.ndarray:
    """v sampled at (coordinate - shift) along one axis with Catmull-Rom, edges clamped."""
    n = v.shape[axis]
    f = np.floor(-shift)
    t = -shift - f
    w = _cr_w(np.array([t]))[0]
    idx = np.arange(n)
    out = np.zeros_like(v, dtype=np.float64)
    for k in range(4):
        src = np.clip(idx + int(f) - 1 + k, 0, n - 1)
        out += w[k] * np.take(v, src, axis=axis)
    return out


def translate(ref: np.ndarray, t) -> np.ndarray:
    """target(x) = ref(x - t) by separable Catmull-Rom (exact tricubic resampling)."""
    v = ref.astype(np.float64)
    v = _shift_axis(v, t[0], 2)
    v = _shift_axis(v, t[1], 1)
    v = _shift_axis(v, t[2], 0)
    top = np.iinfo(ref.dtype).max
    return np.clip(np.rint(v), 0, top).astype(ref.dtype)


def warp(ref: np.ndarray, inverse_map, chunk_slices: int = 8) -> np.ndarray:
    """target(x) = ref(inverse_map(x)) for a vectorised inverse_map((N,3)) -> (N,3)."""
    T, H, W = ref.shape
    out = np.empty(ref.shape, dtype=ref.dtype)
    top = np.iinfo(ref.dtype).max
    yy, xx = np.meshgrid(np.arange(H, dtype=np.float64), np.arange(W, dtype=np.float64), indexing="ij")
    for z0 in range(0, T, chunk_slices):
        zs = np.arange(z0, min(T, z0 + chunk_slices), dtype=np.float64)
        pts = np.stack([np.broadcast_to(xx, (len(zs), H, W)), np.broadcast_to(yy, (len(zs), H, W)),
                        np.broadcast_to(zs[:, None, None], (len(zs), H, W))], axis=-1).reshape(-1, 3)
        v = catmull_rom(ref, inverse_map(pts))
        out[z0:z0 + len(zs)] = np.clip(np.rint(v), 0, top).reshape(len(zs), H, W).astype(ref.dtype)
    return out


def rot_x(deg: float) -> np.ndarray:
    a = np.deg2rad(deg)
    c, s = np.cos(a), np.sin(a)
    return np.array([[1, 0, 0], [0, c, -s], [0, s, c]])


@dataclass
class Affine:
    """phi(X) = c + t + F (X - c)."""
    F: np.ndarray
    t: np.ndarray
    c: np.ndarray

    def forward(self, X):
        return self.c + self.t + (np.asarray(X) - self.c) @ self.F.T

    def inverse(self, x):
        Fi = np.linalg.inv(self.F)
        return self.c + (np.asarray(x) - self.c - self.t) @ Fi.T

    def displacement(self, X):
        return self.forward(X) - np.asarray(X)


def grid_points(n: int, spacing: float, center) -> np.ndarray:
    """n^3 grid, label order x fastest (as a .roi written row by row)."""
    off = (np.arange(n) - (n - 1) / 2.0) * spacing
    z, y, x = np.meshgrid(off + center[2], off + center[1], off + center[0], indexing="ij")
    return np.stack([x.ravel(), y.ravel(), z.ravel()], axis=1)


@dataclass
class Case:
    name: str
    ref: np.ndarray
    tar: np.ndarray
    points: np.ndarray
    truth: np.ndarray                      # imposed displacement at each point
    params: dict = field(default_factory=dict)
    start: np.ndarray | None = None        # starting_point (None: points[0])


def translation_u8(size: int = 128, n: int = 10, spacing: float = 8.0, seed: int = 1,
            t=(1.3, -0.7, 2.1), npts: int = 2000) -> Case:
    """Plain 3-DOF translation case: uint8 size^3, sigma 1.5, cube 31 / 2000 samples,
    ZNSSD, disp_max 5, rigid_trans the nearest integer translation."""
    ref = smoothed_noise((size, size, size), 1.5, np.uint8, seed)
    tar = translate(ref, t)
    c = (size / 2.0,) * 3
    pts = grid_points(n, spacing, c)
    truth = np.tile(np.asarray(t, dtype=np.float64), (len(pts), 1))
    params = dict(ndof=3, obj="znssd", interp="tricubic", geom="cube", subvol_size=31, subvol_npts=npts,
                  disp_max=5, basin_radius=0.0, rigid_trans=tuple(float(np.rint(v)) for v in t))
    return Case("translation-u8", ref, tar, pts, truth, params)


def affine_u16(size: int = 256, n: int = 15, spacing: float = 12.0, seed: int = 2,
            t=(2.4, 0.3, -1.1), rot_deg: float = 1.0, strain_z: float = 0.005, npts: int = 5000,
            dof: int = 12, diameter: int = 41, volumes=None) -> Case:
    """Plain 12-DOF affine case: uint16 size^3, sigma 1.5, rotation about x through the
    volume centre, uniaxial strain along z, sphere 41 / 5000 samples, ZNSSD, disp_max 10,
    processing order starting at the volume centre.  `volumes` = (ref, tar) made by an
    earlier call with the same arguments skips regenerating them."""
    c = np.full(3, (size - 1) / 2.0)
    F = rot_x(rot_deg) @ np.diag([1.0, 1.0, 1.0 + strain_z])
    A = Affine(F, np.asarray(t, dtype=np.float64), c)
    if volumes is not None:
        ref, tar = volumes
        assert ref.shape == (size, size, size) and ref.dtype == np.uint16 and tar.shape == ref.shape
    else:
        ref = smoothed_noise((size, size, size), 1.5, np.uint16, seed)
        tar = warp(ref, A.inverse)
    pts = grid_points(n, spacing, (size / 2.0,) * 3)
    truth = A.displacement(pts)
    start = np.full(3, size / 2.0)
    u0 = A.displacement(start[None])[0]
    params = dict(ndof=dof, obj="znssd", interp="tricubic", geom="sphere", subvol_size=diameter,
                  subvol_npts=npts, disp_max=10, basin_radius=0.0,
                  rigid_trans=tuple(float(np.rint(v)) for v in u0))
    return Case("affine-u16", ref, tar, pts, truth, params, start)


@dataclass
class DeviceCase:
    """A pair generated on the device (dvc_synth_noise / dvc_synth_warp_field): the
    volumes live only in the context's slots."""
    name: str
    shape: tuple
    dtype: type
    points: np.ndarray
    truth: np.ndarray
    params: dict
    start: np.ndarray
    bumps: np.ndarray
    strain: float


def bump_field(X, bumps, strain, center, waves=None):
    """u(X) of dvc_synth_warp_field: Gaussian bumps, plane sinusoids and a uniform
    strain about center."""
    X = np.asarray(X, dtype=np.float64)
    u = strain * (X - np.asarray(center))
    for b in np.asarray(bumps).reshape(-1, 7):
        g = np.exp(-((X - b[:3]) ** 2).sum(axis=-1) / (2.0 * b[3] ** 2))
        u = u + g[..., None] * b[4:7]
    if waves is not None:
        for w in np.asarray(waves).reshape(-1, 7):
            u = u + np.sin(X @ w[:3] + w[3])[..., None] * w[4:7]
    return u


def full_scale_u16(ctx, size: int = 1536, n: int = 100, spacing: float = 14.0, seed: int = 5, nbumps: int = 8,
            amplitude: float = 4.0, width: float = 200.0, volumetric: float = 0.01, npts: int = 8000,
            diameter: int = 41) -> DeviceCase:
    """Full-scale 12-DOF case made on the device: uint16 size^3 smoothed noise (sigma 2),
    target under 8 Gaussian bumps (amplitude 4 voxels, width 200) plus 1 % volumetric
    strain about the centre; n^3 grid at `spacing`, cube 41 / 8000 samples, ZNSSD,
    disp_max 10, processing order from the volume centre."""
    rng = np.random.default_rng(seed)
    shape = (size, size, size)
    ctx.synth_noise(0, shape, np.uint16, 2.0, 1000 + seed)
    dirs = rng.normal(size=(nbumps, 3))
    dirs /= np.linalg.norm(dirs, axis=1, keepdims=True)
    bumps = np.concatenate([rng.uniform(0, size - 1, (nbumps, 3)), np.full((nbumps, 1), width),
                            amplitude * dirs], axis=1)
    strain = (1.0 + volumetric) ** (1.0 / 3.0) - 1.0
    center = np.full(3, (size - 1) / 2.0)
    ctx.synth_warp_field(bumps, strain, center)
    pts = grid_points(n, spacing, (size / 2.0,) * 3)
    truth = bump_field(pts, bumps, strain, center)
    start = np.full(3, size / 2.0)
    u0 = bump_field(start[None], bumps, strain, center)[0]
    params = dict(ndof=12, obj="znssd", interp="tricubic", geom="cube", subvol_size=diameter, subvol_npts=npts,
                  disp_max=10, basin_radius=0.0, rigid_trans=tuple(float(np.rint(v)) for v in u0))
    return DeviceCase("full-u16", shape, np.uint16, pts, truth, params, start, bumps, strain)


def learned_eval_u8(ctx, size: int = 1024, npoints: int = 100_000, seed: int = 4, trans: float = 10.0,
                 amplitude: float = 2.0, wavelength: float = 256.0, margin: float = 48.0) -> DeviceCase:
    """Evaluation pair of the learned proposal, made on the device: uint8 size^3 smoothed
    noise (sigma 2); the target under a uniform translation (each axis up to `trans`
    voxels; a bump far wider than the volume) plus a sinusoidal field (amplitude
    `amplitude` per axis, wavelength `wavelength`, random direction and phase per axis);
    `npoints` points uniform inside `margin`; cube 31 / 2197 samples, ZNSSD, 12-DOF,
    disp_max 12, processing order from the volume centre."""
    rng = np.random.default_rng(seed)
    shape = (size, size, size)
    ctx.synth_noise(0, shape, np.uint8, 2.0, 2000 + seed)
    t = rng.uniform(-trans, trans, 3)
    bumps = np.array([[0.0, 0.0, 0.0, 1e9, *t]])
    waves = []
    for ax in range(3):
        k = rng.normal(size=3)
        k = k / np.linalg.norm(k) * (2 * np.pi / wavelength)
        a = np.zeros(3)
        a[ax] = amplitude
        waves.append(np.concatenate([k, [rng.uniform(0, 2 * np.pi)], a]))
    waves = np.array(waves)
    center = np.full(3, (size - 1) / 2.0)
    ctx.synth_warp_field(bumps, 0.0, center, waves=waves)
    pts = rng.uniform(margin, size - 1 - margin, (npoints, 3))
    truth = bump_field(pts, bumps, 0.0, center, waves)
    start = np.full(3, size / 2.0)
    u0 = bump_field(start[None], bumps, 0.0, center, waves)[0]
    params = dict(ndof=12, obj="znssd", interp="tricubic", geom="cube", subvol_size=31, subvol_npts=2197,
                  disp_max=12, basin_radius=0.0, rigid_trans=tuple(float(np.rint(v)) for v in u0))
    case = DeviceCase("learned-eval-u8", shape, np.uint8, pts, truth, params, start, bumps, 0.0)
    case.waves = waves
    return case


def uneven_u8(ctx, size: int = 512, npoints: int = 50_000, seed: int = 3, amplitude: float = 3.0,
            wavelength: float = 128.0, margin: float = 48.0):
    """Batched uneven workload made on the device: uint8 size^3 smoothed noise (sigma 2),
    the target under three plane sinusoids (amplitude 3 voxels along x, y and z in turn,
    wavelength 128, random directions and phases); `npoints` uniform random points inside
    `margin`; every point its own subvolume: cube or sphere, size uniform in 21..61,
    samples uniform in 1000..10000.  Returns (case, geom, size, nsamp); 6-DOF, disp_max 6,
    processing order from the volume centre."""
    rng = np.random.default_rng(seed)
    shape = (size, size, size)
    ctx.synth_noise(0, shape, np.uint8, 2.0, 3000 + seed)
    waves = []
    for ax in range(3):
        k = rng.normal(size=3)
        k = k / np.linalg.norm(k) * (2 * np.pi / wavelength)
        a = np.zeros(3)
        a[ax] = amplitude
        waves.append(np.concatenate([k, [rng.uniform(0, 2 * np.pi)], a]))
    waves = np.array(waves)
    center = np.full(3, (size - 1) / 2.0)
    ctx.synth_warp_field(np.zeros((0, 7)), 0.0, center, waves=waves)
    pts = rng.uniform(margin, size - 1 - margin, (npoints, 3))
    truth = bump_field(pts, np.zeros((0, 7)), 0.0, center, waves)
    geom = rng.integers(0, 2, npoints).astype(np.int32)
    sub = rng.integers(21, 62, npoints).astype(np.int32)
    nsamp = rng.integers(1000, 10001, npoints).astype(np.int32)
    start = np.full(3, size / 2.0)
    u0 = bump_field(start[None], np.zeros((0, 7)), 0.0, center, waves)[0]
    params = dict(ndof=6, obj="znssd", interp="tricubic", geom="cube", subvol_size=31, subvol_npts=2000,
                  disp_max=6, basin_radius=0.0, rigid_trans=tuple(float(np.rint(v)) for v in u0))
    case = DeviceCase("uneven-u8", shape, np.uint8, pts, truth, params, start, np.zeros((0, 7)), 0.0)
    case.waves = waves
    return case, geom, sub, nsamp
