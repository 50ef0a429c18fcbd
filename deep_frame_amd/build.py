"""Build libdfmi.so (HIP for gfx950) in-tree with hipcc.

Every ``csrc/*.hip`` file is compiled to an object in ``_build/`` (in parallel)
and linked into ``deep_frame_amd/libdfmi.so``.  Packed FP32 instructions are
disabled for all device code.
"""
import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor
from pathlib import Path

HERE = Path(__file__).resolve().parent
ROOT = HERE.parent
SRC = HERE / "csrc"
OUT = HERE / "libdfmi.so"
BUILD = HERE / "_build"
ARCH = os.environ.get("DFMI_ARCH", "gfx950")
FLAGS = ["-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}", "-Xclang", "-target-feature", "-Xclang",
         "-packed-fp32-ops", f"-I{ROOT / 'include'}", f"-I{SRC}", "-Wno-unused-command-line-argument"]


def hipcc():
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", "hipcc"):
        if cand and (Path(cand).exists() or cand == "hipcc"):
            return cand
    raise RuntimeError("hipcc not found")


def build(verbose=False):
    BUILD.mkdir(exist_ok=True)
    cc = hipcc()
    sources = sorted(SRC.glob("*.hip"))

    def compile_one(src):
        obj = BUILD / (src.stem + ".o")
        cmd = [cc, *FLAGS, "-c", str(src), "-o", str(obj)]
        res = subprocess.run(cmd, capture_output=True, text=True)
        if res.returncode != 0:
            raise RuntimeError(f"hipcc failed for {src.name}:\n{res.stderr[-6000:]}")
        if verbose and res.stderr.strip():
            print(res.stderr[-3000:], file=sys.stderr)
        return obj

    with ThreadPoolExecutor(max_workers=min(8, len(sources))) as pool:
        objs = list(pool.map(compile_one, sources))
    tmp = OUT.with_suffix(".so.tmp")
    cmd = [cc, "-shared", f"--offload-arch={ARCH}", "-fPIC", "-Wl,--no-undefined", *map(str, objs), "-o", str(tmp)]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"link failed:\n{res.stderr[-6000:]}")
    os.replace(tmp, OUT)
    return OUT


if __name__ == "__main__":
    print(build(verbose=True))
