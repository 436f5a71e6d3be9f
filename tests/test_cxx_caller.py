"""A C++ caller written against the reference's call surface
(digitalvolumecorrelation_amd/host/dvc_search.h), built with g++ against libdvc_amd.so:

* the reference point loop around Search::process_point (Core/dvc.cpp:233-294), which
  throws Point_Good / Range_Fail / Convg_Fail (Core/Utility.h:77-79) and whose results
  feed the next points' start guesses through the DataCloud records;
* Search::search_pt_setup + Search::obj_val_at(x) and obj_val_at(x, residual)
  (Core/Search.cpp:248-275, 448-488);
* Interpolate::kernels on a headered big-endian uint16 file (Core/Interpolate.cpp:144-258)
  with nearest, tri_lin and tri_cub_Lek;
* the ten obj_* overloads (Core/ObjectiveFunctions.h:31-45);

each against the oracle.  Bar: statuses equal, displacements within 1e-3 voxel,
objectives within 1e-5 relative; interpolated values within 2e-6 of the full range
(FP32 interpolation); objectives of given vectors within 1e-12 relative (FP64 on both
sides)."""
import subprocess
from pathlib import Path

import numpy as np
import pytest

import cli_util as U
from digitalvolumecorrelation_amd import synth

ROOT = Path(__file__).resolve().parent.parent
PKG = ROOT / "digitalvolumecorrelation_amd"
HOST = PKG / "host"
SRC = Path(__file__).resolve().parent / "cpp" / "search_caller.cpp"


def _build(out: Path) -> Path:
    U.ensure_built()
    cmd = ["g++", "-O2", "-std=c++17", "-Wall", f"-I{ROOT / 'include'}", f"-I{HOST}", str(SRC),
           str(HOST / "dvc_io.cpp"), str(HOST / "dvc_search.cpp"), "-o", str(out), f"-L{PKG}", "-ldvc_amd",
           f"-Wl,-rpath,{PKG}", "-pthread"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stdout + r.stderr
    return out


def test_caller_builds(tmp_path):
    # the header surface compiles and links on a machine without a GPU
    exe = _build(tmp_path / "search_caller")
    r = subprocess.run([str(exe)], capture_output=True, text=True)
    assert r.returncode == 1 and "usage" in r.stderr


@pytest.fixture(scope="module")
def caller(tmp_path_factory):
    return _build(tmp_path_factory.mktemp("cxx") / "search_caller")


@pytest.mark.gpu
def test_reference_loop_and_obj_val_at(tmp_path, caller, oracle):
    case = synth.translation_u8(size=64, n=4, spacing=14)     # outer points' boxes leave the volume
    p = dict(case.params, ndof=6)
    start = (32.0, 32.0, 32.0)
    d = U.write_run(tmp_path, case.ref, case.tar, case.points, p, extra={"starting_point": "32.0 32.0 32.0"})
    r = U.run([caller, "loop", "dvc_in"], d, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    lines = r.stdout.splitlines()
    rows = [ln.split() for ln in lines if ln.startswith("pt ")]
    idx = np.array([int(v[1]) for v in rows])
    st = np.array([int(v[2]) for v in rows])
    obj = np.array([float(v[3]) for v in rows])
    par = np.array([[float(x) for x in v[4:16]] for v in rows])

    order, neigh = oracle.sort_order_neighbors(case.points, start=np.array(start))
    run = oracle.make_run(shape=case.ref.shape, bytes_per=1, **p)
    o, _, _ = oracle.dvc_run(run, case.ref, case.tar, case.points, order, neigh, nthreads=8)
    assert np.array_equal(idx, order)
    oo = o[order]
    assert np.array_equal(st, oo["status"])
    assert (st == 0).any() and (st == -1).any()
    good = st == 0
    assert np.abs(par[good, :3] - oo["par"][good, :3]).max() <= 1e-3
    assert (np.abs(obj[good] - oo["obj"][good]) / np.abs(oo["obj"][good])).max() <= 1e-5
    assert np.all(par[~good] == 0.0)

    # obj_val_at at the first point's result, from its neighbours' final records
    n0 = order[0]
    gp = np.zeros((len(case.points), 12))
    gp[idx] = par
    gs = np.full(len(case.points), -3)
    gs[idx] = st
    nb = [m for m in neigh[n0] if gs[m] == 0]
    want_start = np.zeros(3)
    for m in nb:
        want_start += gp[m, :3]
    want_start /= len(nb)
    sl = next(ln.split() for ln in lines if ln.startswith("start "))
    assert np.array_equal(np.array([float(v) for v in sl[1:4]]), want_start)
    ol = next(ln.split() for ln in lines if ln.startswith("objval "))
    a, b, nres, ss = float(ol[1]), float(ol[2]), int(ol[3]), float(ol[4])
    x = gp[n0, :p["ndof"]]
    ov, ores = oracle.obj_val_at(run, case.ref, case.tar, case.points[n0], want_start, x, residual=True)
    print(f"  obj_val_at: library {a:.10e}, oracle {ov:.10e}; residual {nres} samples")
    assert a == b and nres == len(ores)
    assert abs(a - ov) <= 1e-5 * abs(ov)
    assert abs(ss - (ores ** 2).sum()) <= 1e-5 * (ores ** 2).sum()
    assert "far Range_Fail" in lines


@pytest.mark.gpu
@pytest.mark.parametrize("method", [0, 1, 2])
def test_interpolate_kernels_headered_big_endian(tmp_path, caller, oracle, method):
    vol = synth.smoothed_noise((48, 56, 64), 1.5, np.uint16, seed=11)
    header = b"DVC-TEST-HEADER-" + bytes(range(21))
    (tmp_path / "be.raw").write_bytes(header + vol.astype(">u2").tobytes())
    (tmp_path / "le.raw").write_bytes(vol.tobytes())
    rng = np.random.default_rng(5)
    c = np.array([31.3, 27.9, 23.6])
    xyz = c + rng.uniform(-9.0, 9.0, (400, 3))
    np.savetxt(tmp_path / "pts.txt", xyz, fmt="%.17g")
    T, H, W = vol.shape
    out = {}
    for name, endian, hdr in (("be.raw", "big", len(header)), ("le.raw", "little", 0)):
        r = U.run([caller, "interp", name, 2, endian, hdr, W, H, T, "pts.txt", method], tmp_path, timeout=60)
        assert r.returncode == 0, r.stdout + r.stderr
        out[endian] = np.array([float(v) for v in r.stdout.split()])
    assert np.array_equal(out["big"], out["little"])
    # the oracle's interpolator on the same 30-voxel box: a 26-voxel subvolume grown by 2
    run = oracle.make_run(shape=vol.shape, bytes_per=2, subvol_size=26, disp_max=0)
    ctr = np.zeros(3)
    for q in xyz:                      # the caller's sequential centroid
        ctr += q
    ctr /= len(xyz)
    want = oracle.interp(run, vol, ctr, xyz, np.r_[xyz.min(0), xyz.max(0)], method=method)
    assert want is not None
    err = np.abs(out["big"] - want).max()
    print(f"  method {method}: max |library - oracle| {err:.3e}")
    assert err <= 2e-6 * 65535


@pytest.mark.gpu
def test_objective_overloads(tmp_path, caller, oracle):
    rng = np.random.default_rng(7)
    r = rng.random(3001) * 200 + 20
    t = 0.9 * r + 3 + rng.normal(0, 4, len(r))
    np.savetxt(tmp_path / "rt.txt", np.c_[r, t], fmt="%.17g")
    res = U.run([caller, "obj", "rt.txt"], tmp_path, timeout=60)
    assert res.returncode == 0, res.stdout + res.stderr
    lines = res.stdout.splitlines()
    plain = [float(v) for v in lines[0].split()]
    for k, kind in enumerate(("sad", "ssd", "zssd", "nssd", "znssd")):
        want, wres = oracle.objective(kind, r, t, residual=True)
        v, r0 = (float(x) for x in lines[1 + k].split())
        assert abs(plain[k] - want) <= 1e-12 * abs(want), kind
        assert abs(v - want) <= 1e-12 * abs(want), kind
        assert abs(r0 - wres[0]) <= 1e-12 * max(abs(wres[0]), 1e-300) + 1e-15, kind
