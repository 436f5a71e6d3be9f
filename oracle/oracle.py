"""ctypes wrapper of the FP64 CPU oracle (oracle/_build/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, as the reference results to compare against and as the timed CPU restatement of
the reference.  The product (digitalvolumecorrelation_amd) never imports it.
"""
from __future__ import annotations

import ctypes as C
import subprocess
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB = HERE / "_build" / "liboracle.so"

OBJ = {"sad": 0, "ssd": 1, "zssd": 2, "nssd": 3, "znssd": 4}
INTERP = {"nearest": 0, "trilinear": 1, "tricubic": 2}
GEOM = {"cube": 0, "sphere": 1}


class Run(C.Structure):
    _fields_ = [
        ("ndof", C.c_int), ("obj", C.c_int), ("interp", C.c_int), ("geom", C.c_int),
        ("subvol_size", C.c_int), ("subvol_npts", C.c_int), ("disp_max", C.c_int), ("bytes_per", C.c_int),
        ("vol_w", C.c_int), ("vol_h", C.c_int), ("vol_t", C.c_int), ("jacobian", C.c_int),
        ("sphere_seed", C.c_uint), ("pad_", C.c_int), ("basin_radius", C.c_double),
        ("rigid_trans", C.c_double * 3), ("aspect", C.c_double * 3),
    ]


class Result(C.Structure):
    _fields_ = [("status", C.c_int), ("iters", C.c_int), ("obj", C.c_double), ("par", C.c_double * 12)]


RESULT_DTYPE = np.dtype([("status", np.int32), ("iters", np.int32), ("obj", np.float64),
                         ("par", np.float64, (12,))])

_lib = None


def build() -> Path:
    subprocess.run(["make", "-s", "-C", str(HERE)], check=True)
    return LIB


def load() -> C.CDLL:
    global _lib
    if _lib is not None:
        return _lib
    if not LIB.exists():
        build()
    _lib = _bind(C.CDLL(str(LIB)))
    return _lib


def load_planted(define: str) -> C.CDLL:
    """A build of the oracle with one planted mistake (an ORC_PLANT_* macro of
    dvc_oracle.cpp), for the tests that show a check fails on it.  Never the oracle."""
    out = HERE / "_build" / f"liboracle_{define.lower()}.so"
    src = HERE / "dvc_oracle.cpp"
    if not out.exists() or out.stat().st_mtime < src.stat().st_mtime:
        out.parent.mkdir(exist_ok=True)
        subprocess.run(["g++", "-O2", "-std=c++17", "-fPIC", "-fopenmp", "-ffp-contract=off", f"-D{define}",
                        "-shared", "-o", str(out) + ".tmp", str(src)], check=True)
        Path(str(out) + ".tmp").replace(out)
    return _bind(C.CDLL(str(out)))


def _bind(lib: C.CDLL) -> C.CDLL:
    P, I, D, VP = C.POINTER, C.c_int, C.c_double, C.c_void_p
    sig = {
        "orc_version": (I, []),
        "orc_bi_matrix": (None, [P(D)]),
        "orc_constraint_matrix": (None, [P(D)]),
        "orc_subv_num": (I, [I, I]),
        "orc_cloud": (I, [P(Run), P(D), P(D), P(D), P(D)]),
        "orc_sphere_offsets": (I, [C.c_uint, D, I, P(D), P(D)]),
        "orc_affine": (I, [P(Run), P(D), P(D), I, P(D), P(D)]),
        "orc_interp": (I, [P(Run), VP, P(D), P(D), I, P(D), I, P(D), P(D)]),
        "orc_objective": (D, [I, P(D), P(D), I, P(D)]),
        "orc_obj_val_at": (I, [P(Run), VP, VP, P(D), P(D), P(D), I, P(D), P(D)]),
        "orc_sort_order_neighbors": (I, [P(D), I, P(D), P(I), P(I), I]),
        "orc_dvc_run": (I, [P(Run), VP, VP, P(D), I, P(I), P(I), I, I, I, VP, P(D)]),
        "orc_dvc_run_geo": (I, [P(Run), VP, VP, P(D), I, P(I), P(I), I, I, I, P(I), P(I), P(I), VP, P(D)]),
        "orc_dvc_run_start": (I, [P(Run), VP, VP, P(D), I, P(I), P(I), I, I, I, I, P(D), VP, VP, VP, P(I), P(D)]),
        "orc_strain": (I, [P(D), P(D), P(I), P(D), P(I), I, I, I, D, I, I, P(D), P(I), P(D), P(D), P(D), P(D)]),
        "orc_colpiv_solve": (I, [P(D), P(D), I, P(D)]),
        "orc_fullpiv_solve": (I, [P(D), P(D), I, P(D)]),
        "orc_sym3_eigen": (None, [P(D), P(D)]),
    }
    for name, (res, args) in sig.items():
        f = getattr(lib, name)
        f.restype = res
        f.argtypes = args
    return lib


def _dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def _ip(a):
    return a.ctypes.data_as(C.POINTER(C.c_int))


def make_run(*, ndof=3, obj="znssd", interp="tricubic", geom="cube", subvol_size=31, subvol_npts=2000,
             disp_max=5, bytes_per=1, shape=(128, 128, 128), jacobian=0, sphere_seed=1, basin_radius=0.0,
             rigid_trans=(0.0, 0.0, 0.0), aspect=(1.0, 1.0, 1.0)) -> Run:
    r = Run()
    r.ndof, r.obj, r.interp, r.geom = ndof, OBJ.get(obj, obj) if isinstance(obj, str) else obj, \
        INTERP[interp] if isinstance(interp, str) else interp, GEOM[geom] if isinstance(geom, str) else geom
    r.subvol_size, r.subvol_npts, r.disp_max, r.bytes_per = subvol_size, subvol_npts, disp_max, bytes_per
    t, h, w = shape
    r.vol_w, r.vol_h, r.vol_t = w, h, t
    r.jacobian, r.sphere_seed, r.basin_radius = jacobian, sphere_seed, basin_radius
    for q in range(3):
        r.rigid_trans[q] = rigid_trans[q]
        r.aspect[q] = aspect[q]
    return r


def bi_matrix() -> np.ndarray:
    out = np.zeros((64, 64))
    load().orc_bi_matrix(_dp(out))
    return out


def constraint_matrix() -> np.ndarray:
    out = np.zeros((64, 64))
    load().orc_constraint_matrix(_dp(out))
    return out


def subv_num(geom: str, npts: int) -> int:
    return load().orc_subv_num(GEOM[geom], npts)


def cloud(run: Run, pt):
    n = subv_num("cube" if run.geom == 0 else "sphere", run.subvol_npts)
    xyz = np.zeros((n, 3))
    bbox = np.zeros(6)
    ref = np.zeros(3)
    p = np.ascontiguousarray(pt, dtype=np.float64)
    load().orc_cloud(C.byref(run), _dp(p), _dp(xyz), _dp(bbox), _dp(ref))
    return xyz, bbox, ref


def sphere_offsets(seed, rad, num, aspect=(1.0, 1.0, 1.0)):
    out = np.zeros((num, 3))
    a = np.ascontiguousarray(aspect, dtype=np.float64)
    load().orc_sphere_offsets(seed, rad, num, _dp(a), _dp(out))
    return out


def affine(run: Run, pt, par, ndof):
    n = subv_num("cube" if run.geom == 0 else "sphere", run.subvol_npts)
    xyz = np.zeros((n, 3))
    bbox = np.zeros(6)
    p = np.ascontiguousarray(pt, dtype=np.float64)
    pa = np.zeros(12)
    pa[:len(par)] = par
    load().orc_affine(C.byref(run), _dp(p), _dp(pa), ndof, _dp(xyz), _dp(bbox))
    return xyz, bbox


def interp(run: Run, vol: np.ndarray, center, xyz, bbox=None, method=2, grad=False):
    xyz = np.ascontiguousarray(xyz, dtype=np.float64).reshape(-1, 3)
    if bbox is None:
        bbox = np.concatenate([xyz.min(0), xyz.max(0)])
    bbox = np.ascontiguousarray(bbox, dtype=np.float64)
    vol = np.ascontiguousarray(vol)
    out = np.zeros(len(xyz))
    g = np.zeros((len(xyz), 3)) if grad else None
    c = np.ascontiguousarray(center, dtype=np.float64)
    rc = load().orc_interp(C.byref(run), vol.ctypes.data_as(C.c_void_p), _dp(c), _dp(xyz), len(xyz), _dp(bbox),
                           method, _dp(out), _dp(g) if grad else None)
    if rc != 0:
        return None if not grad else (None, None)
    return (out, g) if grad else out


def objective(kind, ref, tar, residual=False):
    r = np.ascontiguousarray(ref, dtype=np.float64)
    t = np.ascontiguousarray(tar, dtype=np.float64)
    res = np.zeros(len(r)) if residual else None
    v = load().orc_objective(OBJ[kind] if isinstance(kind, str) else kind, _dp(r), _dp(t), len(r),
                             _dp(res) if residual else None)
    return (v, res) if residual else v


def obj_val_at(run, ref, tar, point, start, x, residual=False):
    """Search::obj_val_at(x[, residual]) after search_pt_setup for one point (the
    correlated volume's box on point + start); None on Range_Fail."""
    ref = np.ascontiguousarray(ref)
    tar = np.ascontiguousarray(tar)
    p = np.ascontiguousarray(point, dtype=np.float64)
    s = np.ascontiguousarray(start, dtype=np.float64)
    xv = np.ascontiguousarray(x, dtype=np.float64)
    n = load().orc_subv_num(run.geom, run.subvol_npts)
    res = np.zeros(n) if residual else None
    o = C.c_double()
    m = load().orc_obj_val_at(C.byref(run), ref.ctypes.data_as(C.c_void_p), tar.ctypes.data_as(C.c_void_p), _dp(p),
                              _dp(s), _dp(xv), len(xv), C.byref(o), _dp(res) if residual else None)
    if m < 0:
        return None
    return (o.value, res) if residual else o.value


def sort_order_neighbors(points, start=None, nthreads=8):
    p = np.ascontiguousarray(points, dtype=np.float64)
    n = len(p)
    s = np.ascontiguousarray(p[0] if start is None else start, dtype=np.float64)
    k = min(75, n)
    order = np.zeros(n, np.int32)
    neigh = np.zeros((n, k), np.int32)
    load().orc_sort_order_neighbors(_dp(p), n, _dp(s), _ip(order), _ip(neigh), nthreads)
    return order, neigh


def dvc_run(run: Run, ref: np.ndarray, tar: np.ndarray, points, order, neigh, nproc=0, nthreads=1):
    """The serial dvc point loop; returns (results, seconds, points processed)."""
    p = np.ascontiguousarray(points, dtype=np.float64)
    o = np.ascontiguousarray(order, dtype=np.int32)
    nb = np.ascontiguousarray(neigh, dtype=np.int32)
    ref = np.ascontiguousarray(ref)
    tar = np.ascontiguousarray(tar)
    res = np.zeros(len(p), dtype=RESULT_DTYPE)
    sec = C.c_double()
    n = load().orc_dvc_run(C.byref(run), ref.ctypes.data_as(C.c_void_p), tar.ctypes.data_as(C.c_void_p), _dp(p),
                           len(p), _ip(o), _ip(nb), nb.shape[1], nproc, nthreads,
                           res.ctypes.data_as(C.c_void_p), C.byref(sec))
    return res, sec.value, n


def dvc_run_geo(run: Run, ref: np.ndarray, tar: np.ndarray, points, order, neigh, geom, size, npp, nproc=0,
                nthreads=1):
    """The serial loop with per-point subvolume geometry; returns (results, seconds, n)."""
    p = np.ascontiguousarray(points, dtype=np.float64)
    o = np.ascontiguousarray(order, dtype=np.int32)
    nb = np.ascontiguousarray(neigh, dtype=np.int32)
    g = np.ascontiguousarray([GEOM[x] if isinstance(x, str) else int(x) for x in geom], dtype=np.int32)
    sz = np.ascontiguousarray(size, dtype=np.int32)
    m = np.ascontiguousarray(npp, dtype=np.int32)
    ref = np.ascontiguousarray(ref)
    tar = np.ascontiguousarray(tar)
    res = np.zeros(len(p), dtype=RESULT_DTYPE)
    sec = C.c_double()
    n = load().orc_dvc_run_geo(C.byref(run), ref.ctypes.data_as(C.c_void_p), tar.ctypes.data_as(C.c_void_p), _dp(p),
                               len(p), _ip(o), _ip(nb), nb.shape[1], nproc, nthreads, _ip(g), _ip(sz), _ip(m),
                               res.ctypes.data_as(C.c_void_p), C.byref(sec))
    return res, sec.value, n


def dvc_run_start(run: Run, ref: np.ndarray, tar: np.ndarray, points, order, neigh, start, mode=1, nproc=0,
                  nthreads=1, lib=None):
    """Given-start passes (mode 1) or given-start + judge passes (mode 2); returns
    (results, proposal-pass results, judge-pass results, kept, seconds)."""
    p = np.ascontiguousarray(points, dtype=np.float64)
    o = np.ascontiguousarray(order, dtype=np.int32)
    nb = np.ascontiguousarray(neigh, dtype=np.int32)
    st = np.ascontiguousarray(start, dtype=np.float64).reshape(len(p), 12)
    ref = np.ascontiguousarray(ref)
    tar = np.ascontiguousarray(tar)
    res = np.zeros(len(p), dtype=RESULT_DTYPE)
    prop = np.zeros(len(p), dtype=RESULT_DTYPE)
    judge = np.zeros(len(p), dtype=RESULT_DTYPE)
    kept = np.zeros(len(p), dtype=np.int32)
    sec = C.c_double()
    vp = lambda a: a.ctypes.data_as(C.c_void_p)  # noqa: E731
    (lib or load()).orc_dvc_run_start(C.byref(run), vp(ref), vp(tar), _dp(p), len(p), _ip(o), _ip(nb), nb.shape[1],
                                      nproc, nthreads, int(mode), _dp(st), vp(prop), vp(judge), vp(res), _ip(kept),
                                      C.byref(sec))
    return res, prop, judge, kept, sec.value


def strain(pts, disp, status, objmin, neigh, ndp=25, objt=1.0, refill=False, planar=0):
    pts = np.ascontiguousarray(pts, dtype=np.float64)
    disp = np.ascontiguousarray(disp, dtype=np.float64)
    st = np.ascontiguousarray(status, dtype=np.int32)
    om = np.ascontiguousarray(objmin, dtype=np.float64)
    nb = np.ascontiguousarray(neigh, dtype=np.int32)
    n, k = nb.shape
    o = {"disp": np.zeros((n, 3)), "npts": np.zeros(n, np.int32), "swr": np.zeros(n),
         "E": np.zeros((n, 9)), "L": np.zeros((n, 9)), "DG": np.zeros((n, 9))}
    load().orc_strain(_dp(pts), _dp(disp), _ip(st), _dp(om), _ip(nb), k, n, ndp, float(objt), int(refill),
                      int(planar), _dp(o["disp"]), _ip(o["npts"]), _dp(o["swr"]), _dp(o["E"]), _dp(o["L"]),
                      _dp(o["DG"]))
    return o


def colpiv_solve(A, b):
    A = np.ascontiguousarray(A, dtype=np.float64)
    b = np.ascontiguousarray(b, dtype=np.float64)
    x = np.zeros(len(b))
    load().orc_colpiv_solve(_dp(A), _dp(b), len(b), _dp(x))
    return x


def fullpiv_solve(A, b):
    A = np.ascontiguousarray(A, dtype=np.float64)
    b = np.ascontiguousarray(b, dtype=np.float64)
    x = np.zeros(len(b))
    load().orc_fullpiv_solve(_dp(A), _dp(b), len(b), _dp(x))
    return x


def sym3_eigen(S):
    S = np.ascontiguousarray(S, dtype=np.float64).reshape(9)
    lam = np.zeros(3)
    load().orc_sym3_eigen(_dp(S), _dp(lam))
    return lam
