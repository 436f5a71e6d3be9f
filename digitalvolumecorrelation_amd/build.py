"""Build recipe of libdvc_amd.so (hipcc, gfx950) and the host executables.

``python -m digitalvolumecorrelation_amd.build`` compiles everything in place:

* ``digitalvolumecorrelation_amd/libdvc_amd.so`` -- kernels and C ABI (include/dvc_amd.h)
* ``digitalvolumecorrelation_amd/bin/dvc`` and ``bin/strain`` -- the reference CLIs
  (Core/dvc.cpp, Core/strain.cpp) re-hosted on the C ABI
"""
from __future__ import annotations

import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor
from pathlib import Path

PKG = Path(__file__).resolve().parent
ROOT = PKG.parent
CSRC = PKG / "csrc"
HOST = PKG / "host"
LIB = PKG / "libdvc_amd.so"
BIN = PKG / "bin"

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

# Packed FP32 instructions are disabled for all device code; -O3 FP32 code would
# otherwise be lowered to v_pk_* forms.
HIP_FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function", "-Wno-unused-variable", "-Wno-unused-value", "-Wno-unused-result",
]


def _run(cmd: list[str]) -> None:
    print("+", " ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)


# dvc_kernels.hip is compiled as KERNEL_PARTS translation units at once: part 0 holds
# the launch dispatch and the small kernels, parts 1.. the search-kernel instantiations.
KERNEL_PARTS = 6


def build_library() -> Path:
    jobs = []
    for k in range(KERNEL_PARTS):
        jobs.append((CSRC / "dvc_kernels.hip", CSRC / f"dvc_kernels_p{k}.o", [f"-DDVC_PART={k}"]))
    for s in (CSRC / "dvc_api.cpp", CSRC / "dvc_net.hip", CSRC / "dvc_synth.hip"):
        jobs.append((s, CSRC / (s.stem + ".o"), []))

    def compile_one(job):
        src, obj, defs = job
        cmd = [HIPCC, *HIP_FLAGS, *defs, "-x", "hip", "-c", str(src), "-o", str(obj)]
        print("+", " ".join(cmd), flush=True)
        r = subprocess.run(cmd, capture_output=True, text=True)
        return cmd, r

    workers = max(1, min(len(jobs), int(os.environ.get("MAX_JOBS", "0") or 0) or (os.cpu_count() or 1)))
    with ThreadPoolExecutor(max_workers=workers) as ex:
        results = list(ex.map(compile_one, jobs))
    for cmd, r in results:
        msg = "\n".join(l for l in (r.stdout + r.stderr).splitlines() if "packed-fp32-ops" not in l)
        if msg.strip():
            print(msg, flush=True)
        if r.returncode != 0:
            raise subprocess.CalledProcessError(r.returncode, cmd)
    objs = [o for _, o, _ in jobs]
    tmp = LIB.with_suffix(".so.tmp")
    _run([HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", *map(str, objs), "-o", str(tmp)])
    os.replace(tmp, LIB)
    for o in objs:
        o.unlink(missing_ok=True)
    return LIB


def build_host() -> list[Path]:
    """The dvc and strain executables: host C++ over the C ABI (g++, rpath to the lib)."""
    out = []
    if not HOST.exists():
        return out
    BIN.mkdir(exist_ok=True)
    common = [str(HOST / "dvc_io.cpp"), str(HOST / "dvc_search.cpp")]
    for name, main in (("dvc", "dvc_main.cpp"), ("strain", "strain_main.cpp")):
        src = HOST / main
        if not src.exists():
            continue
        exe = BIN / name
        _run(["g++", "-O2", "-std=c++17", "-Wall", f"-I{ROOT / 'include'}", f"-I{HOST}",
              str(src), *common, "-o", str(exe) + ".tmp",
              f"-L{PKG}", "-ldvc_amd", f"-Wl,-rpath,{PKG}", "-Wl,-rpath,$ORIGIN/..",
              "-pthread"])
        os.replace(str(exe) + ".tmp", exe)
        out.append(exe)
    return out


def build_all() -> None:
    build_library()
    build_host()


if __name__ == "__main__":
    build_all()
    sys.exit(0)
