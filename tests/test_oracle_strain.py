"""Pins of the oracle's strain-window fit (Core/strain.cpp:402-591).

A displacement field that the 10-term quadratic model represents exactly must be
recovered exactly: the displacement gradient DG, the fitted displacement, and the
engineering / Lagrangian strains that follow from DG in closed form.
"""
import numpy as np
import pytest

from digitalvolumecorrelation_amd import synth


def _closed_form(G):
    E = 0.5 * (G + G.T)
    L = 0.5 * (G + G.T + G.T @ G)
    pack = lambda S: np.r_[S[0, 0], S[1, 1], S[2, 2], S[0, 1], S[1, 2], S[0, 2], np.sort(np.linalg.eigvalsh(S))[::-1]]
    return pack(E), pack(L)


@pytest.fixture(scope="module")
def grid():
    pts = synth.grid_points(7, 10.0, (60.0, 70.0, 80.0))
    return pts


def test_linear_field_exact(oracle, grid):
    G = np.array([[0.010, 0.002, -0.003], [0.001, -0.004, 0.002], [0.0, 0.003, 0.005]])
    u0 = np.array([0.5, -1.25, 2.0])
    disp = u0 + grid @ G.T
    order, neigh = oracle.sort_order_neighbors(grid)
    st = np.zeros(len(grid), np.int32)
    om = np.full(len(grid), 0.01)
    out = oracle.strain(grid, disp, st, om, neigh, ndp=27)
    assert np.abs(out["DG"] - G.reshape(1, 9)).max() < 1e-9
    assert np.abs(out["disp"] - disp).max() < 1e-8
    E, L = _closed_form(G)
    assert np.abs(out["E"] - E).max() < 1e-9
    assert np.abs(out["L"] - L).max() < 1e-9
    assert np.all(out["npts"] == 27)


def test_quadratic_field_gradient(oracle, grid):
    def field(p):
        x, y, z = p[:, 0], p[:, 1], p[:, 2]
        return np.stack([1e-4 * x * x - 2e-4 * y * z + 0.003 * x, 5e-5 * y * y + 1e-4 * x * z, -1e-4 * z * z + 2e-4 * x * y], 1)

    def grad(p):
        x, y, z = p
        return np.array([[2e-4 * x + 0.003, -2e-4 * z, -2e-4 * y], [1e-4 * z, 1e-4 * y, 1e-4 * x],
                         [2e-4 * y, 2e-4 * x, -2e-4 * z]])

    disp = field(grid)
    order, neigh = oracle.sort_order_neighbors(grid)
    out = oracle.strain(grid, disp, np.zeros(len(grid), np.int32), np.zeros(len(grid)), neigh, ndp=27)
    for i in range(len(grid)):
        assert np.abs(out["DG"][i].reshape(3, 3) - grad(grid[i])).max() < 1e-8


def test_window_filtering(oracle, grid):
    G = np.diag([0.01, 0.02, -0.01])
    disp = grid @ G.T
    order, neigh = oracle.sort_order_neighbors(grid)
    st = np.zeros(len(grid), np.int32)
    om = np.full(len(grid), 0.01)
    st[neigh[0, 1]] = -1                 # a failed neighbour leaves the window
    om[neigh[0, 2]] = 0.5                # above the objmin threshold
    out = oracle.strain(grid, disp, st, om, neigh, ndp=27, objt=0.1)
    assert out["npts"][0] == 25          # strain.cpp:424-434 scans exactly ndp slots
    out = oracle.strain(grid, disp, st, om, neigh, ndp=27, objt=0.1, refill=True)
    assert out["npts"][0] == 27          # -r refills from further neighbours (:435-448)
    assert np.abs(out["DG"][0] - G.reshape(9)).max() < 1e-9
    d_last = np.linalg.norm(grid[0] - grid[neigh[0, 28]])
    assert out["swr"][0] == pytest.approx(d_last, rel=1e-9)


def test_planar_cloud(oracle):
    pts = synth.grid_points(7, 10.0, (60.0, 70.0, 80.0))
    pts = pts[np.isclose(pts[:, 2], pts[0, 2])]            # one z plane -> xy_only
    G = np.array([[0.01, 0.004, 0.0], [0.002, -0.006, 0.0], [0.0, 0.0, 0.0]])
    disp = pts @ G.T
    order, neigh = oracle.sort_order_neighbors(pts)
    out = oracle.strain(pts, disp, np.zeros(len(pts), np.int32), np.zeros(len(pts)), neigh, ndp=25)
    assert np.abs(out["DG"][:, [0, 1, 3, 4]] - G[:2, :2].reshape(4)).max() < 1e-9
    E = 0.5 * (G + G.T)
    lam = np.sort(np.linalg.eigvalsh(E[:2, :2]))[::-1]
    assert np.allclose(out["E"][:, 6:8], lam, atol=1e-12) and np.all(out["E"][:, 8] == 0.0)
