"""Build libcsx.so (HIP for gfx950 + host C++) in-tree with hipcc.

    python -m csim_amd.build          # incremental
    python -m csim_amd.build --force  # rebuild every object
"""
from __future__ import annotations

import argparse
import concurrent.futures as cf
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
SRC = os.path.join(HERE, "csrc")
OBJ = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libcsx.so")
INCLUDE = os.path.join(os.path.dirname(HERE), "include")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"
COMMON = ["-O3", "-fPIC", "-std=c++17", f"-I{INCLUDE}", "-Wall", "-Wno-unused-result",
          "-Wno-unused-function", "-Wno-unused-variable"]
# FP32 packed-math instructions are kept out of every device code object.
DEVICE = [f"--offload-arch={ARCH}", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
          "-munsafe-fp-atomics"]


def sources():
    out = []
    for name in sorted(os.listdir(SRC)):
        if name.endswith((".hip", ".cpp")):
            out.append(os.path.join(SRC, name))
    return out


def _obj_for(src):
    return os.path.join(OBJ, os.path.basename(src) + ".o")


def _deps_newer(src, obj):
    if not os.path.exists(obj):
        return True
    t = os.path.getmtime(obj)
    if os.path.getmtime(src) > t:
        return True
    for name in os.listdir(SRC):
        if name.endswith(".hpp") and os.path.getmtime(os.path.join(SRC, name)) > t:
            return True
    return os.path.getmtime(os.path.join(INCLUDE, "csx.h")) > t


def _compile(src, force):
    obj = _obj_for(src)
    if not force and not _deps_newer(src, obj):
        return obj, None
    cmd = [HIPCC] + COMMON + DEVICE + ["-x", "hip", "-c", src, "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        return obj, f"$ {' '.join(cmd)}\n{r.stdout}\n{r.stderr}"
    return obj, None


def build(force: bool = False, jobs: int | None = None) -> str:
    os.makedirs(OBJ, exist_ok=True)
    srcs = sources()
    jobs = jobs or min(8, os.cpu_count() or 4)
    errors = []
    objs = []
    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        for obj, err in ex.map(lambda s: _compile(s, force), srcs):
            objs.append(obj)
            if err:
                errors.append(err)
    if errors:
        raise RuntimeError("csx build failed:\n" + "\n".join(errors))
    tmp = LIB + ".tmp"
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp] + objs
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"csx link failed:\n$ {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, LIB)
    return LIB


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--force", action="store_true")
    ap.add_argument("-j", type=int, default=None)
    a = ap.parse_args(argv)
    print(build(a.force, a.j))


if __name__ == "__main__":
    sys.exit(main())
