"""Build libpylians_hip.so (gfx950) in-tree with hipcc.

    python -m pylians_amd.build [--jobs N] [--arch gfx950]

Every ``pylians_amd/csrc/*.hip`` file is compiled to an object under
``pylians_amd/_build/`` and linked into ``pylians_amd/libpylians_hip.so``.
"""
import argparse
import concurrent.futures as cf
import glob
import os
import subprocess
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OUT = os.path.join(HERE, "libpylians_hip.so")
BUILD = os.path.join(HERE, "_build")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

FLAGS = [
    "-O3", "-std=c++17", "-fPIC",
    # no packed-FP32 instructions in generated code
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-result",
]


def _compile(src, arch):
    obj = os.path.join(BUILD, os.path.basename(src) + ".o")
    cmd = [HIPCC, "--offload-arch=" + arch] + FLAGS + ["-c", src, "-o", obj]
    t = time.time()
    p = subprocess.run(cmd, capture_output=True, text=True)
    if p.returncode != 0:
        raise RuntimeError("hipcc failed for %s:\n%s\n%s" % (src, " ".join(cmd), p.stderr[-6000:]))
    return obj, time.time() - t


def build(jobs=None, arch="gfx950", verbose=True):
    os.makedirs(BUILD, exist_ok=True)
    srcs = sorted(glob.glob(os.path.join(CSRC, "*.hip")))
    jobs = jobs or max(1, min(len(srcs), os.cpu_count() or 1, 8))
    objs = []
    with cf.ThreadPoolExecutor(jobs) as ex:
        for obj, dt in ex.map(lambda s: _compile(s, arch), srcs):
            objs.append(obj)
            if verbose:
                print("  compiled %-28s %5.1fs" % (os.path.basename(obj), dt), flush=True)
    tmp = OUT + ".tmp"
    cmd = [HIPCC, "--offload-arch=" + arch, "-shared", "-fPIC", "-o", tmp] + objs
    p = subprocess.run(cmd, capture_output=True, text=True)
    if p.returncode != 0:
        raise RuntimeError("link failed:\n%s\n%s" % (" ".join(cmd), p.stderr[-4000:]))
    os.replace(tmp, OUT)
    if verbose:
        print("built", OUT, flush=True)
    return OUT


if __name__ == "__main__":
    ap = argparse.ArgumentParser()
    ap.add_argument("--jobs", type=int, default=None)
    ap.add_argument("--arch", default="gfx950")
    a = ap.parse_args()
    try:
        build(a.jobs, a.arch)
    except RuntimeError as e:
        print(e, file=sys.stderr)
        sys.exit(1)
