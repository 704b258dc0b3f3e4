"""Build libpyfwix_hip.so for gfx950 with hipcc (no rocFFT, no torch extension).

Each ``csrc/*.hip`` file is compiled to an object in ``pyfwix_amd/_build/`` in
parallel, then linked in-tree as ``pyfwix_amd/libpyfwix_hip.so``.  The C++
operator-layer test driver (``tests/cpp/test_operators.cpp`` over
``include/fwix_operators.hpp``) is linked against it as
``tests/cpp/test_operators``.

    python -m pyfwix_amd.build            # build (always recompiles)
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import shutil
import subprocess
import sys
from pathlib import Path

PKG = Path(__file__).resolve().parent
ROOT = PKG.parent
CSRC = PKG / "csrc"
BUILD = PKG / "_build"
LIB = PKG / "libpyfwix_hip.so"
ARCH = os.environ.get("FWIX_OFFLOAD_ARCH", "gfx950")

HIPCC = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"

COMMON = [
    "-O3", "-std=c++17", f"--offload-arch={ARCH}", "-fPIC",
    # packed FP32 (v_pk_*_f32) is kept off: the VALU code is scalar FP32
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-I", str(ROOT / "include"), "-I", str(CSRC),
    "-Wno-unused-command-line-argument",
]


def _compile(src: Path) -> Path:
    obj = BUILD / (src.stem + ".o")
    cmd = [HIPCC, *COMMON, "-c", str(src), "-o", str(obj)]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"hipcc failed on {src.name}:\n{res.stdout}\n{res.stderr}")
    return obj


def build(verbose: bool = True, jobs: int | None = None) -> Path:
    BUILD.mkdir(exist_ok=True)
    srcs = sorted(CSRC.glob("*.hip"))
    jobs = jobs or min(len(srcs), os.cpu_count() or 4, 8)
    with cf.ThreadPoolExecutor(jobs) as ex:
        objs = list(ex.map(_compile, srcs))
    tmp = LIB.with_suffix(".so.tmp")
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", *map(str, objs), "-o", str(tmp)]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"link failed:\n{res.stdout}\n{res.stderr}")
    os.replace(tmp, LIB)
    if verbose:
        print(f"built {LIB} ({LIB.stat().st_size // 1024} KiB) from {len(srcs)} sources")
    build_cpp_tests(verbose)
    build_pyfwix(verbose)
    return LIB


def build_pyfwix(verbose: bool = True) -> Path:
    """The pybind11 module ``pyfwix_amd/_pyFWIX*.so`` (the reference's pyFWIX classes)."""
    import sysconfig

    import pybind11
    src = CSRC / "pyFWIX.cpp"
    out = PKG / ("_pyFWIX" + sysconfig.get_config_var("EXT_SUFFIX"))
    tmp = out.with_name(out.name + ".tmp")
    cmd = [HIPCC, "-O2", "-std=c++17", "-fPIC", "-shared", f"--offload-arch={ARCH}", "-fvisibility=hidden",
           "-I", str(ROOT / "include"), "-I", sysconfig.get_paths()["include"], "-I", pybind11.get_include(),
           str(src), "-L", str(PKG), "-lpyfwix_hip", "-Wl,-rpath,$ORIGIN", "-o", str(tmp),
           "-Wno-unused-command-line-argument"]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"pyFWIX module failed to build:\n{res.stdout}\n{res.stderr}")
    os.replace(tmp, out)
    if verbose:
        print(f"built {out}")
    return out


def build_cpp_tests(verbose: bool = True) -> Path | None:
    src = ROOT / "tests" / "cpp" / "test_operators.cpp"
    if not src.exists():
        return None
    exe = src.with_suffix("")
    cmd = [HIPCC, "-O2", "-std=c++17", f"--offload-arch={ARCH}", "-I", str(ROOT / "include"), str(src),
           "-L", str(PKG), "-lpyfwix_hip", "-Wl,-rpath,$ORIGIN/../../pyfwix_amd", "-o", str(exe),
           "-Wno-unused-command-line-argument"]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"C++ operator test driver failed to build:\n{res.stdout}\n{res.stderr}")
    if verbose:
        print(f"built {exe}")
    return exe


if __name__ == "__main__":
    build()
    sys.exit(0)
