"""Helpers of the dvc / strain program tests: run directories with raw volumes, ROI
and keyword input files in the reference's formats (Core/InputRead.cpp:123-455,
632-716), and readers of the .disp / .sort.csv / strain CSV outputs."""
import subprocess
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent.parent
BIN = ROOT / "digitalvolumecorrelation_amd" / "bin"


def ensure_built():
    if not (BIN / "dvc").exists() or not (BIN / "strain").exists():
        import sys
        sys.path.insert(0, str(ROOT))
        from digitalvolumecorrelation_amd import build
        if not build.LIB.exists():
            build.build_library()
        build.build_host()
    return BIN


def write_run(d: Path, ref: np.ndarray, tar: np.ndarray, points, params: dict, labels=None, header=b"",
              extra: dict | None = None, big_endian=False, omit=()):
    """Raw volumes (x fastest) with an optional header, the ROI file and dvc_in."""
    d.mkdir(parents=True, exist_ok=True)
    for name, vol in (("ref.raw", ref), ("cor.raw", tar)):
        v = vol.astype(vol.dtype.newbyteorder(">")) if big_endian else vol
        (d / name).write_bytes(header + v.tobytes())
    labels = list(range(1, len(points) + 1)) if labels is None else labels
    with open(d / "roi.txt", "w") as f:
        f.write("label\tx\ty\tz\n")
        for lab, p in zip(labels, points):
            f.write(f"{lab}\t{float(p[0])!r}\t{float(p[1])!r}\t{float(p[2])!r}\n")
    T, H, W = ref.shape
    kw = {
        "reference_filename": "ref.raw", "correlate_filename": "cor.raw", "point_cloud_filename": "roi.txt",
        "output_filename": "result",
        "vol_bit_depth": 8 * ref.dtype.itemsize, "vol_endian": "big" if big_endian else "little",
        "vol_hdr_lngth": len(header), "vol_wide": W, "vol_high": H, "vol_tall": T,
        "subvol_geom": params["geom"], "subvol_size": params["subvol_size"], "subvol_npts": params["subvol_npts"],
        "subvol_thresh": "off", "disp_max": params["disp_max"], "num_srch_dof": params["ndof"],
        "obj_function": params["obj"], "interp_type": params["interp"],
        "rigid_trans": " ".join(repr(float(v)) for v in params.get("rigid_trans", (0.0, 0.0, 0.0))),
        "basin_radius": params.get("basin_radius", 0.0),
        "subvol_aspect": "1.0 1.0 1.0", "num_points_to_process": 0,
    }
    kw.update(extra or {})
    lines = ["# dvc run control file written by the tests", ""]
    for k, v in kw.items():
        if k in omit:
            continue
        lines.append(f"{k}\t{v}\t\t### comment")
    (d / "dvc_in").write_text("\n".join(lines) + "\n")
    return d


def run(args, cwd, timeout=600):
    return subprocess.run([str(a) for a in args], cwd=cwd, capture_output=True, text=True, timeout=timeout)


def read_disp(path: Path):
    """Rows of a .disp file: label, x, y, z, status, objmin, u, v, w."""
    lines = path.read_text().splitlines()
    assert lines[0] == "n\tx\ty\tz\tstatus\tobjmin\tu\tv\tw"
    rows = [ln.split("\t") for ln in lines[1:]]
    lab = np.array([int(r[0]) for r in rows])
    xyz = np.array([[float(v) for v in r[1:4]] for r in rows])
    st = np.array([int(r[4]) for r in rows])
    obj = np.array([float(r[5]) for r in rows])
    uvw = np.array([[float(v) for v in r[6:9]] for r in rows])
    return lab, xyz, st, obj, uvw, rows


def read_sort(path: Path):
    return np.array([[int(v) for v in ln.split(",")] for ln in path.read_text().splitlines() if ln.strip()])


def cfmt(v: float) -> str:
    """C++ ostream default formatting of a double (printf %g, 6 significant digits)."""
    return "%g" % v
