"""Build recipe of the native library ``libcvr.so`` (HIP for gfx950 + host C++).

Every ``csrc/*.hip`` and ``csrc/host/*.cpp`` file is compiled by ``hipcc`` into an
object under ``csrc/_build/`` and linked in-tree into
``cpp_volume_rendering_amd/libcvr.so``; ``csrc/app/cvr_cli.cpp`` becomes the headless
application ``cpp_volume_rendering_amd/cvr_cli``.  Run
``python -m cpp_volume_rendering_amd.build``.
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import shutil
import subprocess
import sys
from pathlib import Path

PKG = Path(__file__).resolve().parent
CSRC = PKG / "csrc"
OBJ = CSRC / "_build"
LIB = PKG / "libcvr.so"
CLI = PKG / "cvr_cli"
INCLUDE = PKG.parent / "include"

ARCH = os.environ.get("CVR_OFFLOAD_ARCH", "gfx950")
HIPCC = os.environ.get("HIPCC", shutil.which("hipcc") or "/opt/rocm/bin/hipcc")

COMMON = [
    "-O3", "-std=c++17", "-fPIC", f"-I{INCLUDE}", f"-I{CSRC}",
    "-Wall", "-Wno-unused-variable", "-Wno-unused-function", "-Wno-unused-but-set-variable",
]
# Device code: gfx950 only; the packed-fp32 feature stays off (its host-pass warning is
# harmless).  MFMA accumulators live in VGPRs: the network's ReLU and bf16 packing then
# read them directly instead of copying each one out of an AGPR first.
DEVICE = [f"--offload-arch={ARCH}", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
          "-munsafe-fp-atomics", "-mllvm", "-amdgpu-mfma-vgpr-form"]


def _sources():
    hip = sorted(CSRC.glob("*.hip"))
    cpp = sorted((CSRC / "host").glob("*.cpp"))
    return hip, cpp


def _compile(src: Path, is_hip: bool, verbose: bool) -> Path:
    obj = OBJ / (src.stem + (".hip.o" if is_hip else ".o"))
    cmd = [HIPCC, *COMMON]
    if is_hip:
        cmd += [*DEVICE, "-x", "hip"]
    cmd += ["-c", str(src), "-o", str(obj)]
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    errs = "\n".join(l for l in (r.stderr or "").splitlines() if "packed-fp32-ops" not in l)
    if r.returncode != 0:
        raise RuntimeError(f"compile failed: {src.name}\n{errs}")
    if verbose and errs.strip():
        print(errs, file=sys.stderr)
    return obj


def build(verbose: bool = False, jobs: int | None = None) -> Path:
    OBJ.mkdir(parents=True, exist_ok=True)
    hip, cpp = _sources()
    jobs = jobs or min(8, (os.cpu_count() or 4))
    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        futs = [ex.submit(_compile, s, True, verbose) for s in hip]
        futs += [ex.submit(_compile, s, False, verbose) for s in cpp]
        objs = [f.result() for f in futs]
    tmp = LIB.with_suffix(".so.tmp")
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", str(tmp), *map(str, objs)]
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("link failed:\n" + r.stderr)
    os.replace(tmp, LIB)
    # the headless application, linked against the library next to it
    app = CSRC / "app" / "cvr_cli.cpp"
    cmd = [HIPCC, *COMMON, str(app), "-o", str(CLI) + ".tmp", f"-L{PKG}", "-lcvr", "-Wl,-rpath,$ORIGIN"]
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("cvr_cli build failed:\n" + r.stderr)
    os.replace(str(CLI) + ".tmp", CLI)
    return LIB


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
