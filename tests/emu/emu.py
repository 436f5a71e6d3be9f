"""ctypes wrapper of the CPU build of the kernels (test infrastructure)."""
import ctypes as C
import subprocess
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
ROOT = HERE.parent.parent
LIB = HERE / "_build" / "libdvc_emu.so"


def build(force=False):
    srcs = [HERE / "emu_driver.cpp", HERE / "dvc_emu.h", ROOT / "digitalvolumecorrelation_amd/csrc/dvc_kernels.hip",
            ROOT / "digitalvolumecorrelation_amd/csrc/dvc_geometry.h", ROOT / "digitalvolumecorrelation_amd/csrc/dvc_internal.h"]
    if LIB.exists() and not force and all(LIB.stat().st_mtime > s.stat().st_mtime for s in srcs):
        return LIB
    LIB.parent.mkdir(exist_ok=True)
    subprocess.run(["g++", "-std=c++20", "-O2", "-fPIC", "-shared", "-ffp-contract=off", "-pthread",
                    f"-I{ROOT / 'digitalvolumecorrelation_amd/csrc'}", f"-I{HERE}", "-x", "c++",
                    str(HERE / "emu_driver.cpp"), "-o", str(LIB) + ".tmp"], check=True)
    Path(str(LIB) + ".tmp").replace(LIB)
    return LIB


_lib = None


def load():
    global _lib
    if _lib is None:
        build()
        _lib = C.CDLL(str(LIB))
    return _lib


def search(params, ref, tar, points, order, neigh, nproc=0, threads=64, gather32=True, precompute=False,
           slices=None):
    """slices: None (one workgroup per point), or (max_workgroups, evaluations,
    samples_per_slice) of the slice launches (max_workgroups 0: no limit); then the
    number of sliced chain launches is returned with the results."""
    from digitalvolumecorrelation_amd._lib import RESULT_DTYPE
    lib = load()
    ref = np.ascontiguousarray(ref)
    tar = np.ascontiguousarray(tar)
    T, H, W = ref.shape
    p = np.ascontiguousarray(points, dtype=np.float64)
    o = np.ascontiguousarray(order, dtype=np.int32)
    nb = np.ascontiguousarray(neigh, dtype=np.int32)
    out = np.zeros(len(p), dtype=RESULT_DTYPE)
    sw, se, sp = slices if slices is not None else (-1, 0, 0)
    ns = C.c_int(0)
    rc = lib.emu_search_sliced(C.byref(params), ref.ctypes.data_as(C.c_void_p), tar.ctypes.data_as(C.c_void_p),
                               ref.dtype.itemsize, W, H, T, p.ctypes.data_as(C.c_void_p), len(p),
                               o.ctypes.data_as(C.c_void_p), nproc, nb.ctypes.data_as(C.c_void_p), nb.shape[1], threads,
                               -1 if gather32 else 0, int(precompute), sw, se, sp, out.ctypes.data_as(C.c_void_p),
                               C.byref(ns))
    if rc != 0:
        raise RuntimeError(f"emulated search error flags {rc}")
    return (out, ns.value) if slices is not None else out


def gn_solve(A, b, chol=False, uniform=False):
    """The search kernel's Gauss-Newton solve (gn_step) of A x = b on one emulated wave:
    the column-pivoted QR alone (chol False), or the Cholesky solve with its QR fallback
    (chol True, the search kernel's own choice; A symmetric)."""
    lib = load()
    A = np.ascontiguousarray(A, dtype=np.float64)
    b = np.ascontiguousarray(b, dtype=np.float64)
    x = np.zeros(len(b))
    vp = lambda a: a.ctypes.data_as(C.c_void_p)
    lib.emu_gn_solve.argtypes = [C.c_int, C.c_void_p, C.c_void_p, C.c_void_p, C.c_int]
    if lib.emu_gn_solve(len(b), vp(A), vp(b), vp(x), 2 if uniform else int(chol)) != 0:
        raise ValueError("n must be 3, 6 or 12")
    return x


def strain(pts, disp, status, objmin, neigh, ndp=25, objt=1.0, refill=False, planar=(0, 0, 0)):
    lib = load()
    pts = np.ascontiguousarray(pts, dtype=np.float64)
    disp = np.ascontiguousarray(disp, dtype=np.float64)
    st = np.ascontiguousarray(status, dtype=np.int32)
    om = np.ascontiguousarray(objmin, dtype=np.float64)
    nb = np.ascontiguousarray(neigh, dtype=np.int32)
    n, k = nb.shape
    o = {"disp": np.zeros((n, 3)), "npts": np.zeros(n, np.int32), "swr": np.zeros(n),
         "E": np.zeros((n, 9)), "L": np.zeros((n, 9)), "DG": np.zeros((n, 9))}
    vp = lambda a: a.ctypes.data_as(C.c_void_p)
    lib.emu_strain.argtypes = [C.c_void_p] * 5 + [C.c_int] * 3 + [C.c_double] + [C.c_int] * 4 + [C.c_void_p] * 6
    lib.emu_strain(vp(pts), vp(disp), vp(st), vp(om), vp(nb), k, n, ndp, float(objt), int(refill), *planar,
                   vp(o["disp"]), vp(o["npts"]), vp(o["swr"]), vp(o["E"]), vp(o["L"]), vp(o["DG"]))
    return o


def schedule(order, neigh, cap=0):
    """The chain's launches: dependency levels (cap 0) or launches of at most cap points;
    a list of point-index arrays."""
    lib = load()
    o = np.ascontiguousarray(order, dtype=np.int32)
    nb = np.ascontiguousarray(neigh, dtype=np.int32)
    n = len(o)
    lp = np.zeros(n, np.int32)
    lo = np.zeros(n + 1, np.int32)
    vp = lambda a: a.ctypes.data_as(C.c_void_p)
    lib.emu_schedule.argtypes = [C.c_void_p, C.c_int, C.c_void_p, C.c_int, C.c_int, C.c_void_p, C.c_void_p]
    d = lib.emu_schedule(vp(o), n, vp(nb), nb.shape[1], int(cap), vp(lp), vp(lo))
    return [lp[lo[t]:lo[t + 1]].copy() for t in range(d)]


def spatial_order(points, order, xcds=8):
    """Workgroup order of the reference kernel: the processing position of each block."""
    lib = load()
    p = np.ascontiguousarray(points, dtype=np.float64)
    o = np.ascontiguousarray(order, dtype=np.int32)
    out = np.zeros(len(o), np.int32)
    vp = lambda a: a.ctypes.data_as(C.c_void_p)
    lib.emu_spatial_order.argtypes = [C.c_void_p, C.c_void_p, C.c_int, C.c_int, C.c_void_p]
    lib.emu_spatial_order(vp(p), vp(o), len(o), int(xcds), vp(out))
    return out

