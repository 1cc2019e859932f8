"""Build the native parts in-tree: libcusmc.so (front end + gfx950 kernels +
C ABI), the `cusmc` command-line tool, and the test oracle.

Plain hipcc / g++ invocations, objects compiled in parallel; no CMake.
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import shutil
import subprocess
import sys
from pathlib import Path

ROOT = Path(__file__).resolve().parent.parent
PKG = ROOT / "cusmc_amd"
CSRC = PKG / "csrc"
BUILD = PKG / "_build"
LIB = PKG / "lib"
ORACLE = ROOT / "oracle"

ARCH = os.environ.get("SMC_OFFLOAD_ARCH", "gfx950")
HIPCC = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
CXX = shutil.which("g++") or "g++"

COMMON = ["-std=c++17", "-O3", "-fPIC", "-ffp-contract=off", f"-I{ROOT / 'include'}", f"-I{CSRC}"]
HIP_FLAGS = [f"--offload-arch={ARCH}", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
             "-Wno-unused-command-line-argument"]

HOST_SOURCES = ["network.cpp", "uppaal.cpp", "compile.cpp"]
HIP_SOURCES = ["capi.cpp", "is.cpp", "sim_kernel.hip", "is_kernels.hip", "noalloc.hip"]
# kernels compiled once per (component slots per lane, register policy)
VARIANT_SOURCES = ["sim_variant.hip", "is_variant.hip"]
VARIANTS = {"sim_variant.hip": [(c, o) for c in (1, 2, 4, 8, 16) for o in (0, 3)],
            # the advance kernel holds at most 8 slots per lane (networks of up
            # to 512 components); 16 slots would triple the build time
            # registers unrestricted only: at 168 VGPRs (three waves per
            # SIMD) the advance kernel spills 100-300 bytes per lane
            "is_variant.hip": [(c, 0) for c in (1, 2, 4, 8)]}
# The kernel variants are one long loop each whose expression interpreter,
# log2 polynomial and Philox rounds use dozens of constants; machine-level
# loop-invariant code motion hoists their materialisation out of the loop
# into some fifty VGPRs held for the kernel's lifetime.  Without it the run
# kernel needs 150-190 VGPRs instead of 230-290 (no spills at two waves per
# SIMD, three for up to four slots per lane); IR-level LICM still runs.
VARIANT_FLAGS = ["-mllvm", "-disable-machine-licm"]


def _deps(src: Path, seen: set | None = None) -> set:
    """The source and the local headers it includes, transitively."""
    seen = set() if seen is None else seen
    if src in seen or not src.exists():
        return seen
    seen.add(src)
    for line in src.read_text(errors="replace").splitlines():
        line = line.strip()
        if line.startswith("#include \""):
            name = line.split('"')[1]
            for d in (src.parent, CSRC, ROOT / "include"):
                if (d / name).exists():
                    _deps(d / name, seen)
                    break
    return seen


def _fresh(obj: Path, src: Path) -> bool:
    """SMC_BUILD_INCREMENTAL=1: an object newer than its source and the
    headers it includes is kept (a development shortcut; build() always
    compiles)."""
    if os.environ.get("SMC_BUILD_INCREMENTAL") != "1" or not obj.exists():
        return False
    return obj.stat().st_mtime > max(f.stat().st_mtime for f in _deps(src))


def _run(cmd: list[str]) -> None:
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    out = "\n".join(l for l in r.stdout.splitlines() if "packed-fp32-ops" not in l and "not a recognized feature" not in l)
    if r.returncode != 0:
        raise RuntimeError(f"build step failed ({r.returncode}): {' '.join(cmd)}\n{out}")


def _obj(src: str) -> Path:
    return BUILD / (src.replace(".", "_") + ".o")


def _variant_obj(src: str, c: int, o: int) -> Path:
    return BUILD / f"{src.replace('.', '_')}_c{c}_o{o}.o"


def build_lib() -> Path:
    BUILD.mkdir(parents=True, exist_ok=True)
    LIB.mkdir(parents=True, exist_ok=True)
    jobs, objs = [], []
    for s in VARIANT_SOURCES:  # the longest jobs first
        for c, o in sorted(VARIANTS[s], key=lambda v: -v[0]):
            obj = _variant_obj(s, c, o)
            objs.append(obj)
            if _fresh(obj, CSRC / s):
                continue
            jobs.append([HIPCC, *COMMON, *HIP_FLAGS, *VARIANT_FLAGS, "-x", "hip", f"-DSMC_C={c}", f"-DSMC_OCC={o}", "-c",
                         str(CSRC / s), "-o", str(obj)])
    for s in HIP_SOURCES:
        extra = ["-x", "hip"] if s.endswith(".hip") else []
        objs.append(_obj(s))
        if not _fresh(_obj(s), CSRC / s):
            jobs.append([HIPCC, *COMMON, *HIP_FLAGS, *extra, "-c", str(CSRC / s), "-o", str(_obj(s))])
    for s in HOST_SOURCES:
        objs.append(_obj(s))
        if not _fresh(_obj(s), CSRC / s):
            jobs.append([CXX, *COMMON, "-Wall", "-c", str(CSRC / s), "-o", str(_obj(s))])
    workers = max(1, min(len(jobs) or 1, int(os.environ.get("MAX_JOBS", "0")) or os.cpu_count() or 8))
    with cf.ThreadPoolExecutor(max_workers=workers) as ex:
        list(ex.map(_run, jobs))
    out = LIB / "libcusmc.so"
    tmp = LIB / "libcusmc.so.tmp"
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-Wl,--no-undefined", "-o", str(tmp),
          *[str(o) for o in objs]])
    os.replace(tmp, out)
    return out


def build_cli() -> Path:
    out = LIB / "cusmc"
    tmp = LIB / "cusmc.tmp"
    _run([CXX, *COMMON, "-Wall", str(CSRC / "cli_main.cpp"), "-o", str(tmp),
          f"-L{LIB}", "-lcusmc", "-Wl,-rpath,$ORIGIN"])
    os.replace(tmp, out)
    return out


def build_oracle() -> Path:
    out = ORACLE / "liboracle.so"
    tmp = ORACLE / "liboracle.so.tmp"
    _run([CXX, *COMMON, "-Wall", "-shared", "-pthread", str(ORACLE / "oracle.cpp"), "-o", str(tmp)])
    os.replace(tmp, out)
    return out


def build_all() -> None:
    with cf.ThreadPoolExecutor(max_workers=2) as ex:
        f_lib = ex.submit(build_lib)
        f_or = ex.submit(build_oracle)
        f_lib.result()
        f_or.result()
    if (CSRC / "cli_main.cpp").exists():
        build_cli()


if __name__ == "__main__":
    what = sys.argv[1:] or ["all"]
    for w in what:
        {"all": build_all, "lib": build_lib, "cli": build_cli, "oracle": build_oracle}[w]()
    print("built", ", ".join(what))
