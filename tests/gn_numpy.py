"""An independent numpy restatement of the per-point Gauss-Newton search, for pinning the
oracle's (test infrastructure only).

Follows Core/Search.cpp:194-246 (min_Lev_Mar: at most 20 iterations; stop when
i > 0 and |obj - obj_old| < 1e-6 or |update[0:3]| < 0.01, the update being the previous
step; the step solves J^T J dx = -J^T e), with the analytic Jacobian of the residual,
the warp of Core/SearchParams.cpp:53-107 and Core/FloatingCloud.cpp:148-286
(x' = c + u + (I + S) A (x - c)) and the objectives of Core/ObjectiveFunctions.cpp
(SSD: e = t - r; ZNSSD: e = (t - mean t)/|t - mean t| - (r - mean r)/|r - mean r|,
obj = sum e^2 / 4).  Interpolation is separable Catmull-Rom (the interior form of the
Lekien-Marsden tricubic of Core/Interpolate.cpp), written here from its weights; the
cases keep every sample's 4x4x4 support inside the interpolation box.
"""
import numpy as np


def cr_weights(t):
    t2, t3 = t * t, t * t * t
    w = np.stack([0.5 * (-t3 + 2 * t2 - t), 0.5 * (3 * t3 - 5 * t2 + 2), 0.5 * (-3 * t3 + 4 * t2 + t),
                  0.5 * (t3 - t2)], -1)
    d = np.stack([0.5 * (-3 * t2 + 4 * t - 1), 0.5 * (9 * t2 - 10 * t), 0.5 * (-9 * t2 + 8 * t + 1),
                  0.5 * (3 * t2 - 2 * t)], -1)
    return w, d


def cr_sample(vol, xyz):
    """Values and gradients of the Catmull-Rom interpolant of vol[z, y, x] at xyz [n, 3]."""
    v = vol.astype(np.float64)
    i = np.floor(xyz).astype(np.int64)
    f = xyz - i
    wx, dx = cr_weights(f[:, 0])
    wy, dy = cr_weights(f[:, 1])
    wz, dz = cr_weights(f[:, 2])
    val = np.zeros(len(xyz))
    g = np.zeros((len(xyz), 3))
    for c in range(4):
        zz = i[:, 2] - 1 + c
        for b in range(4):
            yy = i[:, 1] - 1 + b
            for a in range(4):
                xx = i[:, 0] - 1 + a
                s = v[zz, yy, xx]
                val += wx[:, a] * wy[:, b] * wz[:, c] * s
                g[:, 0] += dx[:, a] * wy[:, b] * wz[:, c] * s
                g[:, 1] += wx[:, a] * dy[:, b] * wz[:, c] * s
                g[:, 2] += wx[:, a] * wy[:, b] * dz[:, c] * s
    return val, g


def rotation(phi, the, psi):
    cf, sf, ct, st, cp, sp = np.cos(phi), np.sin(phi), np.cos(the), np.sin(the), np.cos(psi), np.sin(psi)
    A = np.array([[ct * cf, ct * sf, -st],
                  [sp * st * cf - cp * sf, sp * st * sf + cp * cf, ct * sp],
                  [cp * st * cf + sp * sf, cp * st * sf - sp * cf, ct * cp]])
    dphi = np.array([[-ct * sf, ct * cf, 0.0],
                     [-sp * st * sf - cp * cf, sp * st * cf - cp * sf, 0.0],
                     [-cp * st * sf + sp * cf, cp * st * cf + sp * sf, 0.0]])
    dthe = np.array([[-st * cf, -st * sf, -ct],
                     [sp * ct * cf, sp * ct * sf, -st * sp],
                     [cp * ct * cf, cp * ct * sf, -st * cp]])
    dpsi = np.array([[0.0, 0.0, 0.0],
                     [cp * st * cf + sp * sf, cp * st * sf - sp * cf, ct * cp],
                     [-sp * st * cf + cp * sf, -sp * st * sf - cp * cf, -ct * sp]])
    return A, (dphi, dthe, dpsi)


def warp(stable, c, par, ndof):
    """Warped sample positions [n, 3] and their derivatives [n, 3, ndof]."""
    rel = stable - c
    u = par[:3]
    D = np.zeros((len(stable), 3, ndof))
    D[:, 0, 0] = D[:, 1, 1] = D[:, 2, 2] = 1.0
    if ndof == 3:
        return stable + u, D
    A, dA = rotation(*par[3:6])
    S = np.zeros((3, 3))
    if ndof == 12:
        e = par[6:12]
        S = np.array([[e[0], e[3], e[5]], [e[3], e[1], e[4]], [e[5], e[4], e[2]]])
    IS = np.eye(3) + S
    w = rel @ A.T
    x = c + u + w @ IS.T
    for q in range(3):
        D[:, :, 3 + q] = rel @ (IS @ dA[q]).T
    if ndof == 12:
        D[:, 0, 6] = w[:, 0]
        D[:, 1, 7] = w[:, 1]
        D[:, 2, 8] = w[:, 2]
        D[:, 0, 9], D[:, 1, 9] = w[:, 1], w[:, 0]
        D[:, 1, 10], D[:, 2, 10] = w[:, 2], w[:, 1]
        D[:, 0, 11], D[:, 2, 11] = w[:, 2], w[:, 0]
    return x, D


def residual(kind, r, t, G):
    """Residual e, Jacobian J (rows: samples) and objective of SSD or ZNSSD."""
    if kind == "ssd":
        e = t - r
        return e, G, float(e @ e)
    tc, rc = t - t.mean(), r - r.mean()
    N, R = np.sqrt(tc @ tc), np.sqrt(rc @ rc)
    e = tc / N - rc / R
    Gc = G - G.mean(0)
    J = Gc / N - np.outer(tc, tc @ G) / N ** 3
    return e, J, float(e @ e) / 4.0


def search(ref, tar, point, offsets, start, ndof, kind):
    """One point's Gauss-Newton search from a given start; (par, obj, iterations)."""
    stable = point + offsets
    r, _ = cr_sample(ref, stable)
    par = np.array(start[:ndof], dtype=np.float64)
    upd = np.zeros(ndof)
    obj_old = 0.0
    iters = 0
    for i in range(20):
        iters += 1
        x, D = warp(stable, point, par, ndof)
        t, g = cr_sample(tar, x)
        G = np.einsum("nk,nkp->np", g, D)
        e, J, obj = residual(kind, r, t, G)
        if i > 0 and (abs(obj - obj_old) < 1e-6 or np.sqrt(upd[:3] @ upd[:3]) < 0.01):
            return par, obj, iters
        upd = np.linalg.solve(J.T @ J, -(J.T @ e))
        par = par + upd
        obj_old = obj
    x, D = warp(stable, point, par, ndof)
    t, _ = cr_sample(tar, x)
    return par, residual(kind, r, t, np.zeros((len(t), ndof)))[2], iters
