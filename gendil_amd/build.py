"""Build recipe for libgendil_amd.so (gfx950 HIP kernels + C-ABI), in-tree.

Every source under gendil_amd/csrc is compiled by hipcc for gfx950 into
_build/ (in parallel) and linked into gendil_amd/libgendil_amd.so.
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
BUILD = ROOT / "_build"
LIB = PKG / "libgendil_amd.so"

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"
FLAGS = [
    "-O3",
    "-std=c++20",
    f"--offload-arch={ARCH}",
    "-fPIC",
    "-Wall",
    "-Wno-unused-function",
    # keep hipcc from emitting packed FP32 instructions
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    f"-I{ROOT / 'include'}",
    f"-I{CSRC}",
]


def _sources():
    return sorted(list(CSRC.glob("*.hip")) + list(CSRC.glob("*.cpp")))


def _compile(src: Path) -> tuple[Path, str]:
    obj = BUILD / (src.name + ".o")
    cmd = [HIPCC, *FLAGS, "-c", str(src), "-o", str(obj)]
    if src.suffix == ".cpp":
        cmd[1:1] = ["-x", "hip"]
    res = subprocess.run(cmd, capture_output=True, text=True)
    msgs = "\n".join(
        line for line in (res.stdout + res.stderr).splitlines() if "packed-fp32-ops" not in line
    )
    if res.returncode != 0:
        raise RuntimeError(f"hipcc failed for {src.name}:\n{msgs}")
    return obj, msgs


def _up_to_date(src: Path) -> bool:
    """The object exists and is newer than its source and every header."""
    obj = BUILD / (src.name + ".o")
    if not obj.exists():
        return False
    headers = list(CSRC.glob("*.hpp")) + list((ROOT / "include").rglob("*.h*"))
    newest = max([src.stat().st_mtime] + [h.stat().st_mtime for h in headers])
    return obj.stat().st_mtime > newest


def build(jobs: int | None = None, verbose: bool = False, incremental: bool = False) -> Path:
    """Compile all HIP/C++ sources for gfx950 and link the shared library
    (with `incremental`, only sources whose object is out of date)."""
    BUILD.mkdir(exist_ok=True)
    srcs = _sources()
    jobs = jobs or min(len(srcs), max(1, (os.cpu_count() or 4)))
    todo = [s for s in srcs if not (incremental and _up_to_date(s))]
    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        fresh = dict(ex.map(_compile, todo))
    results = [(BUILD / (s.name + ".o"), fresh.get(BUILD / (s.name + ".o"), "")) for s in srcs]
    for obj, msgs in results:
        if verbose and msgs.strip():
            print(msgs, file=sys.stderr)
    tmp = LIB.with_suffix(".so.tmp")
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-o", str(tmp)] + [str(o) for o, _ in results]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"link failed:\n{res.stdout}{res.stderr}")
    shutil.move(str(tmp), str(LIB))
    return LIB


if __name__ == "__main__":
    print(build(verbose=True, incremental="--incremental" in sys.argv))
