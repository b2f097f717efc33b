"""Build libisg.so (the MI355X solver) in-tree with hipcc for gfx950.

    python inverse-solver-gpu_amd/build.py [--jobs N]

Each translation unit is compiled to an object in csrc/_build/ in parallel,
then linked into inverse-solver-gpu_amd/libisg.so.  HIP sources are compiled
with packed-FP32 instructions disabled (-packed-fp32-ops).
"""
import argparse
import concurrent.futures as cf
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OBJ = os.path.join(CSRC, "_build")
LIB = os.path.join(HERE, "libisg.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"

COMMON = ["-O3", "-std=c++17", "-fPIC", "-Wall", "-Wno-unused-result", "-I" + CSRC,
          "-I" + os.path.join(HERE, "..", "include")]
HIPFLAGS = ["-x", "hip", "--offload-arch=" + ARCH, "-Xclang", "-target-feature", "-Xclang",
            "-packed-fp32-ops", "-munsafe-fp-atomics"]

# (source, is_hip)
SOURCES = [
    ("isg_plan.cpp", False),
    ("isg_phantom.cpp", False),
    ("isg_io.cpp", False),
    ("isg_engine.hip", True),
    ("isg_dbim.hip", True),
    ("isg_warmnet.hip", True),
    ("isg_capi.cpp", True),
]


def _compile(src, is_hip):
    out = os.path.join(OBJ, os.path.splitext(src)[0] + ".o")
    cmd = [HIPCC] + COMMON + (HIPFLAGS if is_hip else []) + ["-c", os.path.join(CSRC, src), "-o", out]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("compile failed: %s\n%s\n%s" % (" ".join(cmd), r.stdout, r.stderr))
    return out


def build(jobs=4, verbose=False):
    os.makedirs(OBJ, exist_ok=True)
    srcs = [(s, h) for s, h in SOURCES if os.path.exists(os.path.join(CSRC, s))]
    with cf.ThreadPoolExecutor(max_workers=max(1, jobs)) as ex:
        objs = list(ex.map(lambda a: _compile(*a), srcs))
    cmd = [HIPCC, "--offload-arch=" + ARCH, "-shared", "-fPIC", "-o", LIB + ".tmp"] + objs + ["-lrocblas"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("link failed: %s\n%s" % (r.stdout, r.stderr))
    os.replace(LIB + ".tmp", LIB)
    if verbose:
        print("built", LIB)
    return LIB


if __name__ == "__main__":
    ap = argparse.ArgumentParser()
    ap.add_argument("--jobs", type=int, default=4)
    a = ap.parse_args()
    try:
        build(a.jobs, verbose=True)
    except RuntimeError as e:
        print(e, file=sys.stderr)
        sys.exit(1)
