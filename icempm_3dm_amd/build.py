"""Builds the gfx950 shared library ``icempm_3dm_amd/lib/libicempm_3dm.so``.

Every ``csrc/*.hip`` and ``csrc/*.cpp`` file is compiled by hipcc for gfx950
(in parallel), then linked into one shared library exporting the C-ABI of
``include/icempm_3dm.h``. Run ``python -m icempm_3dm_amd.build`` or call
:func:`build`.
"""
from __future__ import annotations

import concurrent.futures
import os
import shutil
import subprocess
import sys
from pathlib import Path

PKG = Path(__file__).resolve().parent
ROOT = PKG.parent
CSRC = PKG / "csrc"
OBJ = PKG / "_build"
LIB_DIR = PKG / "lib"
LIB = LIB_DIR / "libicempm_3dm.so"
ARCH = "gfx950"

HIPCC_FLAGS = [
    f"--offload-arch={ARCH}",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-munsafe-fp-atomics",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall",
    f"-I{ROOT / 'include'}",
    f"-I{CSRC}",
]


def hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and Path(cand).exists():
            return cand
    raise RuntimeError("hipcc not found: the HIP toolchain is required to build icempm_3dm_amd")


def _compile(src: Path, extra: list[str]) -> Path:
    obj = OBJ / (src.name + ".o")
    lang = ["-x", "hip"] if src.suffix == ".hip" else ["-x", "c++"]
    cmd = [hipcc(), *HIPCC_FLAGS, *extra, *lang, "-c", str(src), "-o", str(obj)]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"compile failed: {' '.join(cmd)}\n{res.stdout}\n{res.stderr}")
    if res.stderr.strip():
        # surface warnings, except hipcc's host-pass note on the packed-fp32 feature
        lines = [l for l in res.stderr.splitlines() if "packed-fp32-ops" not in l]
        if any(l.strip() for l in lines):
            sys.stderr.write("\n".join(lines) + "\n")
    return obj


def build(extra_flags: list[str] | None = None, jobs: int | None = None) -> Path:
    """Compiles and links the library; returns its path."""
    extra = list(extra_flags or [])
    OBJ.mkdir(exist_ok=True)
    LIB_DIR.mkdir(exist_ok=True)
    sources = sorted(CSRC.glob("*.hip")) + sorted(CSRC.glob("*.cpp"))
    if not sources:
        raise RuntimeError(f"no sources under {CSRC}")
    jobs = jobs or min(len(sources), max(1, (os.cpu_count() or 2)))
    with concurrent.futures.ThreadPoolExecutor(max_workers=jobs) as ex:
        objs = list(ex.map(lambda s: _compile(s, extra), sources))
    tmp = LIB.with_suffix(".so.tmp")
    cmd = [hipcc(), f"--offload-arch={ARCH}", "-shared", "-fPIC", *map(str, objs), "-lz", "-o", str(tmp)]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"link failed: {' '.join(cmd)}\n{res.stdout}\n{res.stderr}")
    os.replace(tmp, LIB)
    return LIB


CPP_DRIVERS = {"model3d_cycles": ROOT / "tests" / "cpp" / "model3d_cycles.cpp"}
CPP_BIN = ROOT / "tests" / "cpp" / "bin"


def build_cpp_drivers() -> list[Path]:
    """Host programs written against the C++ mirrors of the reference's
    classes (include/*.hpp), linked to the library: tests run them as a
    C++ caller of the reference would."""
    CPP_BIN.mkdir(parents=True, exist_ok=True)
    out = []
    for name, src in CPP_DRIVERS.items():
        exe = CPP_BIN / name
        tmp = exe.with_suffix(".tmp")
        cmd = ["g++", "-std=c++17", "-O2", "-Wall", "-Werror", f"-I{ROOT / 'include'}", str(src), f"-L{LIB_DIR}",
               "-licempm_3dm", "-Wl,-rpath,$ORIGIN/../../../icempm_3dm_amd/lib", "-o", str(tmp)]
        res = subprocess.run(cmd, capture_output=True, text=True)
        if res.returncode != 0:
            raise RuntimeError(f"build failed: {' '.join(cmd)}\n{res.stdout}\n{res.stderr}")
        os.replace(tmp, exe)
        out.append(exe)
    return out


def build_oracle() -> Path:
    """Builds the CPU oracle library (test infrastructure) with its Makefile."""
    res = subprocess.run(["make", "-s", "-C", str(ROOT / "oracle")], capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"oracle build failed:\n{res.stdout}\n{res.stderr}")
    return ROOT / "oracle" / "lib" / "libmpm_oracle.so"


if __name__ == "__main__":
    print(build())
