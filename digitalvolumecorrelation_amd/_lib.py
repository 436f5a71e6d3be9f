"""ctypes binding of libdvc_amd.so (include/dvc_amd.h).

The library is the product path: every call below runs HIP kernels on a gfx950
device.  There is no CPU fallback -- a missing library or device raises.
"""
from __future__ import annotations

import ctypes as C
import os
from pathlib import Path

import numpy as np

# DVC_AMD_LIB selects another build of the library (same ABI), e.g. one compiled with
# different tuning macros.
LIB_PATH = Path(os.environ.get("DVC_AMD_LIB", Path(__file__).resolve().parent / "libdvc_amd.so"))

POINT_GOOD, RANGE_FAIL, CONVG_FAIL, NOT_SEARCH = 0, -1, -2, -3
# start modes of a prepared search (dvc_set_start_mode)
START_CHAIN, START_GIVEN, START_JUDGED = 0, 1, 2
OBJ = {"sad": 0, "ssd": 1, "zssd": 2, "nssd": 3, "znssd": 4}
INTERP = {"nearest": 0, "trilinear": 1, "tricubic": 2}
GEOM = {"cube": 0, "sphere": 1}

ERRORS = {0: "ok", -1: "invalid argument", -2: "HIP failure", -3: "bad call order",
          -5: "no gfx950 device", -6: "outside the interpolation box"}
ERR_RANGE = -6


class RunParams(C.Structure):
    """dvc_run_params: the RunControl fields of Core/Utility.h:90-134 the path reads."""
    _fields_ = [
        ("ndof", C.c_int32), ("obj", C.c_int32), ("interp", C.c_int32), ("geom", C.c_int32),
        ("subvol_size", C.c_int32), ("subvol_npts", C.c_int32), ("disp_max", C.c_int32),
        ("sphere_seed", C.c_uint32), ("basin_radius", C.c_double),
        ("rigid_trans", C.c_double * 3), ("aspect", C.c_double * 3),
    ]


class PointResult(C.Structure):
    """dvc_point_result: ResultRecord (Core/DataCloud.h:45-56) plus iterations."""
    _fields_ = [("status", C.c_int32), ("iters", C.c_int32), ("obj", C.c_double),
                ("par", C.c_double * 12)]


RESULT_DTYPE = np.dtype([("status", np.int32), ("iters", np.int32), ("obj", np.float64),
                         ("par", np.float64, (12,))])
assert RESULT_DTYPE.itemsize == C.sizeof(PointResult)

_lib = None


def load() -> C.CDLL:
    """Load libdvc_amd.so; raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not LIB_PATH.exists():
        raise RuntimeError(f"{LIB_PATH} is missing: run `python -m digitalvolumecorrelation_amd.build`")
    lib = C.CDLL(str(LIB_PATH))
    P, I, D, VP = C.POINTER, C.c_int32, C.c_double, C.c_void_p
    sig = {
        "dvc_create": (I, [P(VP), C.c_int]),
        "dvc_destroy": (None, [VP]),
        "dvc_last_error": (C.c_char_p, [VP]),
        "dvc_set_stream": (I, [VP, VP]),
        "dvc_last_kernel_ms": (I, [VP, P(C.c_float)]),
        "dvc_volume_upload": (I, [VP, C.c_int, VP, C.c_int64, C.c_int, C.c_int, C.c_int64, C.c_int64, C.c_int64]),
        "dvc_volume_attach": (I, [VP, C.c_int, VP, C.c_int, C.c_int64, C.c_int64, C.c_int64]),
        "dvc_volume_download": (I, [VP, C.c_int, VP]),
        "dvc_interpolate": (I, [VP, C.c_int, P(RunParams), P(D), P(D), I, P(D), C.c_int, P(D), P(D)]),
        "dvc_interpolate_box": (I, [VP, C.c_int, P(D), P(D), P(D), P(D), I, P(D), C.c_int, P(D), P(D)]),
        "dvc_objective": (I, [VP, C.c_int, P(D), P(D), I, P(D), P(D)]),
        "dvc_obj_val_at": (I, [VP, P(RunParams), P(D), P(D), P(D), I, P(D), P(D), I, P(I)]),
        "dvc_obj_val_at_noalloc": (I, [VP, P(RunParams), P(D), P(D), P(D), I, P(D), P(D), I, P(I)]),
        "dvc_search_points": (I, [VP, P(RunParams), P(D), I, P(I), I, P(I), I, VP]),
        "dvc_search_prepare": (I, [VP, P(RunParams), P(D), I, P(I), I, P(I), I]),
        "dvc_search_execute": (I, [VP]),
        "dvc_search_fetch": (I, [VP, VP]),
        "dvc_set_block_threads": (I, [VP, C.c_int]),
        "dvc_set_round_capacity": (I, [VP, I]),
        "dvc_set_launch_mode": (I, [VP, C.c_int]),
        "dvc_search_info": (I, [VP, C.c_void_p, C.c_void_p, C.c_void_p]),
        "dvc_set_slice_launches": (I, [VP, I, I, I]),
        "dvc_search_set_strain": (I, [VP, I, D, I, I]),
        "dvc_search_fetch_strain": (I, [VP, VP, VP, VP, VP, VP, VP]),
        "dvc_search_slice_info": (I, [VP, C.c_void_p, C.c_void_p, C.c_void_p]),
        "dvc_set_start_mode": (I, [VP, C.c_int]),
        "dvc_set_point_geometry": (I, [VP, P(I), P(I), P(I), I]),
        "dvc_search_autotune": (I, [VP, C.c_int, P(I), P(C.c_float)]),
        "dvc_search_set_starts": (I, [VP, VP, C.c_int]),
        "dvc_search_fetch_passes": (I, [VP, VP, VP, VP]),
        "dvc_strain_fit": (I, [VP, P(D), P(D), P(I), P(D), P(I), I, I, I, D, I, I,
                               P(D), P(I), P(D), P(D), P(D), P(D)]),
        "dvc_net_create": (I, [VP, C.c_int, C.c_int, P(VP)]),
        "dvc_net_destroy": (None, [VP]),
        "dvc_net_last_error": (C.c_char_p, [VP]),
        "dvc_net_num_params": (C.c_int64, []),
        "dvc_net_params": (I, [VP, P(C.c_float), C.c_int64, C.c_int]),
        "dvc_net_crop": (I, [VP, VP, VP, I, VP, VP, VP]),
        "dvc_device_alloc": (I, [VP, P(VP), C.c_int64]),
        "dvc_device_free": (I, [VP, VP]),
        "dvc_device_copy": (I, [VP, VP, VP, C.c_int64, C.c_int]),
        "dvc_synchronize": (I, [VP]),
        "dvc_net_forward": (I, [VP, VP, I, VP, VP]),
        "dvc_net_gradients": (I, [VP, VP, VP, I, P(C.c_float), P(C.c_float)]),
        "dvc_net_train_step": (I, [VP, VP, VP, I, C.c_float, P(C.c_float)]),
        "dvc_search_propose": (I, [VP, VP]),
        "dvc_synth_warp_tiles": (I, [VP, C.c_int, P(D), C.c_int64]),
        "dvc_synth_noise": (I, [VP, C.c_int, C.c_int, C.c_int64, C.c_int64, C.c_int64, D, C.c_uint64]),
        "dvc_synth_warp_field": (I, [VP, P(D), I, P(D), I, D, P(D), I]),
        "dvc_sort_order_neighbors": (I, [P(D), I, P(D), P(I), P(I), I]),
        "dvc_knn_order_neighbors": (I, [P(D), I, P(D), I, P(I), P(I), I]),
    }
    for name, (res, args) in sig.items():
        f = getattr(lib, name)
        f.restype = res
        f.argtypes = args
    _lib = lib
    return lib


def exported_symbols() -> list[str]:
    """Entry points declared in include/dvc_amd.h."""
    hdr = Path(__file__).resolve().parent.parent / "include" / "dvc_amd.h"
    import re
    return sorted(set(re.findall(r"^\s*(?:int64_t|int|void|const char)\s*\*?\s*(dvc_\w+)\s*\(", hdr.read_text(), re.M)))


def _dp(a: np.ndarray):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def _ip(a: np.ndarray):
    return a.ctypes.data_as(C.POINTER(C.c_int32))


def make_params(*, ndof=3, obj="znssd", interp="tricubic", geom="cube", subvol_size=31,
                subvol_npts=2000, disp_max=5, sphere_seed=1, basin_radius=0.0,
                rigid_trans=(0.0, 0.0, 0.0), aspect=(1.0, 1.0, 1.0)) -> RunParams:
    p = RunParams()
    p.ndof = ndof
    p.obj = OBJ[obj] if isinstance(obj, str) else int(obj)
    p.interp = INTERP[interp] if isinstance(interp, str) else int(interp)
    p.geom = GEOM[geom] if isinstance(geom, str) else int(geom)
    p.subvol_size = subvol_size
    p.subvol_npts = subvol_npts
    p.disp_max = disp_max
    p.sphere_seed = sphere_seed
    p.basin_radius = basin_radius
    for q in range(3):
        p.rigid_trans[q] = rigid_trans[q]
        p.aspect[q] = aspect[q]
    return p


class DvcError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"{ERRORS.get(code, code)}: {msg}")
        self.code = code


class Context:
    """One device context (dvc_create / dvc_destroy)."""

    def __init__(self, device: int = 0):
        self.lib = load()
        h = C.c_void_p()
        rc = self.lib.dvc_create(C.byref(h), device)
        if rc != 0:
            raise DvcError(rc, "dvc_create failed (a gfx950 GPU is required)")
        self.h = h
        self._deps = []            # weak references to networks and device arrays of this context

    def _adopt(self, obj):
        import weakref
        self._deps.append(weakref.ref(obj))

    def close(self):
        # networks and device arrays first: their destructors queue on this context
        for r in getattr(self, "_deps", []):
            o = r()
            if o is not None:
                o.close() if hasattr(o, "close") else o.free()
        self._deps = []
        if getattr(self, "h", None):
            self.lib.dvc_destroy(self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, rc: int, allow=()):
        if rc != 0 and rc not in allow:
            raise DvcError(rc, self.lib.dvc_last_error(self.h).decode())
        return rc

    def set_stream(self, stream_ptr: int | None):
        """Run the context's work on this hipStream_t (None: its own stream)."""
        self._check(self.lib.dvc_set_stream(self.h, C.c_void_p(stream_ptr) if stream_ptr else None))

    def synchronize(self):
        self._check(self.lib.dvc_synchronize(self.h))

    # ---- volumes
    def upload(self, slot: int, vol: np.ndarray, big_endian: bool = False, header: bytes = b""):
        vol = np.ascontiguousarray(vol)
        if vol.dtype not in (np.uint8, np.uint16):
            raise ValueError("volumes are uint8 or uint16")
        t, h, w = vol.shape
        raw = np.frombuffer(header + vol.tobytes(), dtype=np.uint8) if header else vol.view(np.uint8).reshape(-1)
        self._check(self.lib.dvc_volume_upload(self.h, slot, raw.ctypes.data_as(C.c_void_p), len(header),
                                               vol.dtype.itemsize, int(big_endian), w, h, t))

    def download(self, slot: int, shape, dtype) -> np.ndarray:
        """The volume in `slot` (shape (tall, high, wide)) copied back to the host."""
        out = np.empty(shape, dtype)
        self._check(self.lib.dvc_volume_download(self.h, slot, out.ctypes.data_as(C.c_void_p)))
        return out

    def warp_tiles(self, tile: int, maps: np.ndarray):
        """dvc_synth_warp_tiles: slot 1 = slot 0 warped tile by tile (maps [ntiles, 12])."""
        m = np.ascontiguousarray(maps, dtype=np.float64)
        self._check(self.lib.dvc_synth_warp_tiles(self.h, int(tile), _dp(m), len(m)))

    def synth_noise(self, slot: int, shape, dtype, sigma: float, seed: int):
        """dvc_synth_noise: smoothed-noise volume of shape (tall, high, wide) made on the device."""
        t, h, w = shape
        self._check(self.lib.dvc_synth_noise(self.h, slot, np.dtype(dtype).itemsize, w, h, t, float(sigma), int(seed)))

    def synth_warp_field(self, bumps: np.ndarray, strain: float, center, iterations: int = 6, waves=None):
        """dvc_synth_warp_field: slot 1 = slot 0 under Gaussian bumps, plane sinusoids and
        a uniform strain."""
        b = np.ascontiguousarray(bumps, dtype=np.float64).reshape(-1, 7)
        w = np.ascontiguousarray(np.zeros((0, 7)) if waves is None else waves, dtype=np.float64).reshape(-1, 7)
        c = np.ascontiguousarray(center, dtype=np.float64)
        self._check(self.lib.dvc_synth_warp_field(self.h, _dp(b), len(b), _dp(w), len(w), float(strain), _dp(c),
                                                  int(iterations)))

    def attach(self, slot: int, dev_ptr: int, bytes_per: int, shape):
        t, h, w = shape
        self._check(self.lib.dvc_volume_attach(self.h, slot, C.c_void_p(dev_ptr), bytes_per, w, h, t))

    # ---- interpolation / objective
    def interpolate(self, slot, params, center, xyz, bbox, method=2, grad=False):
        xyz = np.ascontiguousarray(xyz, dtype=np.float64).reshape(-1, 3)
        n = xyz.shape[0]
        out = np.zeros(n)
        g = np.zeros((n, 3)) if grad else None
        c = np.ascontiguousarray(center, dtype=np.float64)
        b = np.ascontiguousarray(bbox, dtype=np.float64)
        rc = self.lib.dvc_interpolate(self.h, slot, C.byref(params), _dp(c), _dp(xyz), n, _dp(b), method,
                                      _dp(out), _dp(g) if grad else None)
        self._check(rc, allow=(ERR_RANGE,))
        if rc == ERR_RANGE:
            return None if not grad else (None, None)
        return (out, g) if grad else out

    def objective(self, kind, ref, tar, residual=False):
        r = np.ascontiguousarray(ref, dtype=np.float64)
        t = np.ascontiguousarray(tar, dtype=np.float64)
        o = C.c_double()
        res = np.zeros(len(r)) if residual else None
        self._check(self.lib.dvc_objective(self.h, OBJ[kind] if isinstance(kind, str) else kind, _dp(r), _dp(t),
                                           len(r), C.byref(o), _dp(res) if residual else None))
        return (o.value, res) if residual else o.value

    def obj_val_at(self, params, point, x, start=None, residual=False):
        """dvc_obj_val_at: Search::obj_val_at(x[, residual]) for one point whose
        correlated-volume box sits on point + start (default: params' rigid_trans).
        Returns the objective, or (objective, residual)."""
        p = np.ascontiguousarray(point, dtype=np.float64)
        xv = np.ascontiguousarray(x, dtype=np.float64)
        st = None if start is None else np.ascontiguousarray(start, dtype=np.float64)
        n = C.c_int32(0)
        o = C.c_double()
        self._check(self.lib.dvc_obj_val_at(self.h, C.byref(params), _dp(p), _dp(st) if st is not None else None,
                                            _dp(xv), len(xv), C.byref(o), None, 0, C.byref(n)))
        if not residual:
            return o.value
        res = np.zeros(n.value)
        self._check(self.lib.dvc_obj_val_at(self.h, C.byref(params), _dp(p), _dp(st) if st is not None else None,
                                            _dp(xv), len(xv), C.byref(o), _dp(res), len(res), C.byref(n)))
        return o.value, res

    # ---- search
    def set_block_threads(self, n: int):
        self._check(self.lib.dvc_set_block_threads(self.h, n))

    def set_launch_mode(self, graph: bool):
        self._check(self.lib.dvc_set_launch_mode(self.h, int(graph)))

    def prepare(self, params, points, order, neigh, nproc=0):
        self._pts = np.ascontiguousarray(points, dtype=np.float64)
        self._ord = np.ascontiguousarray(order, dtype=np.int32)
        self._nb = np.ascontiguousarray(neigh, dtype=np.int32)
        self._npts = self._pts.shape[0]
        self._check(self.lib.dvc_search_prepare(self.h, C.byref(params), _dp(self._pts), self._npts, _ip(self._ord),
                                                nproc, _ip(self._nb), self._nb.shape[1]))

    def autotune(self, reps: int = 3):
        """Pick the fastest threads-per-point for the prepared search; (threads, ms)."""
        bt, ms = C.c_int32(), C.c_float()
        self._check(self.lib.dvc_search_autotune(self.h, reps, C.byref(bt), C.byref(ms)))
        return bt.value, ms.value

    def set_slice_launches(self, max_workgroups: int = 0, evaluations: int = 0, samples_per_workgroup: int = 0):
        """Slice launches of the chain (one launch per evaluation, a point's samples split
        over workgroups) for launches of at most max_workgroups workgroups (0 automatic,
        -1 never); `evaluations` slice launches before the one-workgroup search resumes."""
        self._check(self.lib.dvc_set_slice_launches(self.h, int(max_workgroups), int(evaluations),
                                                    int(samples_per_workgroup)))

    def set_search_strain(self, ndp=27, objt=1.0, refill=False, planar=0):
        """Fit the strain window of every execute's results on the device (ndp 0: off)."""
        self._check(self.lib.dvc_search_set_strain(self.h, int(ndp), float(objt), int(bool(refill)), int(planar)))

    def fetch_search_strain(self):
        """Outputs of the last execute's strain fit, as Context.strain returns them."""
        n = self._npts
        o = {"disp": np.zeros((n, 3)), "npts": np.zeros(n, np.int32), "swr": np.zeros(n),
             "E": np.zeros((n, 9)), "L": np.zeros((n, 9)), "DG": np.zeros((n, 9))}
        vp = lambda a: a.ctypes.data_as(C.c_void_p)  # noqa: E731
        self._check(self.lib.dvc_search_fetch_strain(self.h, vp(o["disp"]), vp(o["npts"]), vp(o["swr"]), vp(o["E"]),
                                                     vp(o["L"]), vp(o["DG"])))
        return o

    def slice_info(self):
        """(sliced chain launches, workgroups per point, slice launches per chain launch)."""
        v = [C.c_int32(), C.c_int32(), C.c_int32()]
        self._check(self.lib.dvc_search_slice_info(self.h, *[C.byref(x) for x in v]))
        return tuple(x.value for x in v)

    def set_round_capacity(self, points: int):
        """Points per chain launch for the next prepare: 0 automatic (one round of the
        chip), -1 dependency levels only, else that many."""
        self._check(self.lib.dvc_set_round_capacity(self.h, int(points)))

    def set_point_geometry(self, geom=None, size=None, nsamp=None):
        """Per-point subvolume shape (0 cube, 1 sphere or names), size and sample count for
        the next prepare; no arguments: one geometry for all points."""
        if geom is None:
            self._check(self.lib.dvc_set_point_geometry(self.h, None, None, None, 0))
            return
        g = np.ascontiguousarray([GEOM[x] if isinstance(x, str) else int(x) for x in geom], dtype=np.int32)
        s = np.ascontiguousarray(size, dtype=np.int32)
        m = np.ascontiguousarray(nsamp, dtype=np.int32)
        if not (len(g) == len(s) == len(m)):
            raise ValueError("geometry arrays differ in length")
        self._keep_pg = (g, s, m)
        self._check(self.lib.dvc_set_point_geometry(self.h, _ip(g), _ip(s), _ip(m), len(g)))

    def set_start_mode(self, mode: int):
        """START_CHAIN (the reference's serial start guess), START_GIVEN or START_JUDGED;
        applied by the next prepare."""
        self._check(self.lib.dvc_set_start_mode(self.h, int(mode)))

    def set_starts(self, start):
        """Start warps [npts, 12] of a prepared given-start search: a host array, or a
        device pointer (int) to npts*12 float64 values."""
        if isinstance(start, int):
            self._check(self.lib.dvc_search_set_starts(self.h, C.c_void_p(start), 1))
            return
        st = np.ascontiguousarray(start, dtype=np.float64).reshape(self._npts, 12)
        self._check(self.lib.dvc_search_set_starts(self.h, st.ctypes.data_as(C.c_void_p), 0))

    def fetch_passes(self):
        """(proposal-pass results, judge-pass results, kept) of a START_JUDGED execute."""
        a = np.zeros(self._npts, dtype=RESULT_DTYPE)
        b = np.zeros(self._npts, dtype=RESULT_DTYPE)
        kept = np.zeros(self._npts, dtype=np.int32)
        self._check(self.lib.dvc_search_fetch_passes(self.h, a.ctypes.data_as(C.c_void_p),
                                                     b.ctypes.data_as(C.c_void_p), kept.ctypes.data_as(C.c_void_p)))
        return a, b, kept

    def _info(self):
        v = [C.c_int32(), C.c_int32(), C.c_int32()]
        self._check(self.lib.dvc_search_info(self.h, *[C.byref(x) for x in v]))
        return [x.value for x in v]

    def levels(self) -> int:
        """Chain launches of the prepared search: dependency levels, or launches of at most
        the round capacity when a level is wider."""
        return self._info()[0]

    def block_threads(self) -> int:
        return self._info()[1]

    def execute(self):
        self._check(self.lib.dvc_search_execute(self.h))

    def fetch(self) -> np.ndarray:
        res = np.zeros(self._npts, dtype=RESULT_DTYPE)
        self._check(self.lib.dvc_search_fetch(self.h, res.ctypes.data_as(C.c_void_p)))
        return res

    def search(self, params, points, order, neigh, nproc=0) -> np.ndarray:
        """The serial point loop (dvc_search_points), whatever the start mode."""
        self._pts = np.ascontiguousarray(points, dtype=np.float64)
        self._ord = np.ascontiguousarray(order, dtype=np.int32)
        self._nb = np.ascontiguousarray(neigh, dtype=np.int32)
        self._npts = self._pts.shape[0]
        res = np.zeros(self._npts, dtype=RESULT_DTYPE)
        self._check(self.lib.dvc_search_points(self.h, C.byref(params), _dp(self._pts), self._npts, _ip(self._ord),
                                               nproc, _ip(self._nb), self._nb.shape[1],
                                               res.ctypes.data_as(C.c_void_p)))
        return res

    def kernel_ms(self) -> float:
        ms = C.c_float()
        self._check(self.lib.dvc_last_kernel_ms(self.h, C.byref(ms)))
        return ms.value

    # ---- strain
    def strain(self, pts, disp, status, objmin, neigh, ndp=25, objt=1.0, refill=False, planar=0):
        pts = np.ascontiguousarray(pts, dtype=np.float64)
        disp = np.ascontiguousarray(disp, dtype=np.float64)
        st = np.ascontiguousarray(status, dtype=np.int32)
        om = np.ascontiguousarray(objmin, dtype=np.float64)
        nb = np.ascontiguousarray(neigh, dtype=np.int32)
        n, k = nb.shape
        o = {"disp": np.zeros((n, 3)), "npts": np.zeros(n, np.int32), "swr": np.zeros(n),
             "E": np.zeros((n, 9)), "L": np.zeros((n, 9)), "DG": np.zeros((n, 9))}
        self._check(self.lib.dvc_strain_fit(self.h, _dp(pts), _dp(disp), _ip(st), _dp(om), _ip(nb), k, n, ndp,
                                            float(objt), int(refill), int(planar), _dp(o["disp"]), _ip(o["npts"]),
                                            _dp(o["swr"]), _dp(o["E"]), _dp(o["L"]), _dp(o["DG"])))
        return o


class DeviceArray:
    """A device allocation of the context (dvc_device_alloc), shaped like a numpy
    array.  ``ptr`` is its address; ``row(i)`` the address of row i of the first axis."""

    def __init__(self, ctx: Context, shape, dtype):
        self.ctx = ctx
        self.shape = tuple(int(v) for v in (shape if isinstance(shape, (tuple, list)) else (shape,)))
        self.dtype = np.dtype(dtype)
        self.nbytes = int(np.prod(self.shape)) * self.dtype.itemsize
        p = C.c_void_p()
        ctx._check(ctx.lib.dvc_device_alloc(ctx.h, C.byref(p), self.nbytes))
        self.ptr = int(p.value or 0)
        ctx._adopt(self)

    @classmethod
    def from_host(cls, ctx: Context, arr: np.ndarray) -> "DeviceArray":
        a = np.ascontiguousarray(arr)
        d = cls(ctx, a.shape, a.dtype)
        d.upload(a)
        return d

    def row(self, i: int) -> int:
        return self.ptr + int(i) * (self.nbytes // self.shape[0])

    def upload(self, arr: np.ndarray, row: int = 0):
        a = np.ascontiguousarray(arr, dtype=self.dtype)
        self.ctx._check(self.ctx.lib.dvc_device_copy(self.ctx.h, C.c_void_p(self.row(row)),
                                                     a.ctypes.data_as(C.c_void_p), a.nbytes, 0))

    def download(self) -> np.ndarray:
        out = np.empty(self.shape, self.dtype)
        self.ctx._check(self.ctx.lib.dvc_device_copy(self.ctx.h, out.ctypes.data_as(C.c_void_p),
                                                     C.c_void_p(self.ptr), self.nbytes, 1))
        return out

    def free(self):
        if getattr(self, "ptr", 0) and getattr(self.ctx, "h", None):
            self.ctx.lib.dvc_device_free(self.ctx.h, C.c_void_p(self.ptr))
        self.ptr = 0

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def knn_order_neighbors(points: np.ndarray, start=None, k: int = 75, nthreads: int = 0):
    """Order and k nearest neighbours through a grid (ties in index order)."""
    lib = load()
    p = np.ascontiguousarray(points, dtype=np.float64)
    n = p.shape[0]
    s = np.ascontiguousarray(p[0] if start is None else start, dtype=np.float64)
    k = min(k, n)
    order = np.zeros(n, np.int32)
    neigh = np.zeros((n, k), np.int32)
    rc = lib.dvc_knn_order_neighbors(_dp(p), n, _dp(s), k, _ip(order), _ip(neigh), nthreads)
    if rc != k:
        raise DvcError(rc, "knn_order_neighbors failed")
    return order, neigh


def sort_order_neighbors(points: np.ndarray, start=None, nthreads: int = 0):
    """DataCloud::sort_order_neighbors on the host (no device needed)."""
    lib = load()
    p = np.ascontiguousarray(points, dtype=np.float64)
    n = p.shape[0]
    s = np.ascontiguousarray(p[0] if start is None else start, dtype=np.float64)
    k = min(75, n)
    order = np.zeros(n, np.int32)
    neigh = np.zeros((n, k), np.int32)
    rc = lib.dvc_sort_order_neighbors(_dp(p), n, _dp(s), _ip(order), _ip(neigh), nthreads)
    if rc != k:
        raise DvcError(rc, "sort_order_neighbors failed")
    return order, neigh
