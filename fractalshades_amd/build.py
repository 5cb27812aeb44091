"""Build libfsgpu.so (gfx950) in-tree with hipcc.

The library is compiled into ``fractalshades_amd/_lib/`` so that it travels
with the source tree.  Translation units are compiled in parallel and linked
into one shared object.  Flags:

* ``-ffp-contract=off``: the FP64 path rounds exactly like the reference
  restatement (no fused multiply-add);
* ``-Xclang -target-feature -Xclang -packed-fp32-ops``: no packed FP32 ops.
"""
import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
LIBDIR = os.path.join(HERE, "_lib")
LIB = os.path.join(LIBDIR, "libfsgpu.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"

SOURCES = ["fsg_kernels.hip", "fsg_orbit.hip", "fsg_lds.hip", "fsg_api.hip", "fsg_fp32x.hip", "fsg_serve.hip", "fsg_post.hip",
           "fsg_fixed.cpp"]
# the binary32 form may fuse multiply-adds; the binary64 drop-in may not
CONTRACT = {"fsg_fp32x.hip": "-ffp-contract=fast"}
FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC",
    "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-I", os.path.join(ROOT, "include"), "-I", CSRC,
]


def _run(cmd):
    proc = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    # hipcc's host pass does not know the packed-fp32 feature: drop that noise
    out = "\n".join(
        ln for ln in proc.stdout.splitlines()
        if "packed-fp32-ops" not in ln and "is not a recognized feature" not in ln
    )
    if proc.returncode != 0:
        raise RuntimeError(f"command failed ({proc.returncode}): {' '.join(cmd)}\n{out}")
    return out


def build(verbose=False):
    """Compile every HIP source and link libfsgpu.so.  Returns the path."""
    os.makedirs(LIBDIR, exist_ok=True)
    objs = []
    cmds = []
    for src in SOURCES:
        obj = os.path.join(LIBDIR, os.path.splitext(src)[0] + ".o")
        objs.append(obj)
        flags = [CONTRACT.get(src, f) if f == "-ffp-contract=off" else f for f in FLAGS]
        cmds.append([HIPCC, *flags, "-c", os.path.join(CSRC, src), "-o", obj])
    with ThreadPoolExecutor(max_workers=min(len(cmds), os.cpu_count() or 1)) as ex:
        for out in ex.map(_run, cmds):
            if verbose and out.strip():
                print(out)
    tmp = LIB + ".tmp"
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *objs])
    os.replace(tmp, LIB)
    for obj in objs:
        os.remove(obj)
    if verbose:
        print("built", LIB)
    return LIB


if __name__ == "__main__":
    build(verbose=True)
    sys.exit(0)
