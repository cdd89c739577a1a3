"""Build recipe of libvfxsim3d.so: hipcc for gfx950, in-tree.

The library links against the HIP runtime that PyTorch ships (torch/lib/libamdhip64.so), so a
process that imported torch shares one HIP runtime, one device context and torch's streams with
the library.  Objects go to neural-vfx_amd/_build/, the library to neural-vfx_amd/lib/.
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import subprocess
import sys

PKG_ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))   # neural-vfx_amd
CSRC = os.path.join(PKG_ROOT, "csrc")
BUILD_DIR = os.path.join(PKG_ROOT, "_build")
LIB_DIR = os.path.join(PKG_ROOT, "lib")
LIB_PATH = os.path.join(LIB_DIR, "libvfxsim3d.so")
SOURCES = ("advect.hip", "forces.hip", "pressure.hip", "cg.hip", "capi.hip", "unet.hip", "batch.hip", "fused.hip",
           "hires.hip", "guide.hip", "render.hip", "noalloc.hip")
ARCH = "gfx950"

HIPCC_FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-fPIC", "-std=c++17",
    "-mcode-object-version=5",
    # the parity contract: no contraction of a*b+c into fma, so FP32 results follow the
    # reference's operation order
    "-ffp-contract=off",
    # no packed-FP32 instructions (they have given silent wrong results on these boxes)
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-result",
]


def _torch_lib_dir() -> str:
    import importlib.util
    spec = importlib.util.find_spec("torch")
    if spec is None or spec.origin is None:
        raise RuntimeError("PyTorch is required to build libvfxsim3d.so (its HIP runtime is linked)")
    return os.path.join(os.path.dirname(spec.origin), "lib")


def _hipcc() -> str:
    for c in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", "hipcc"):
        if c and (os.path.sep not in c or os.path.exists(c)):
            return c
    return "hipcc"


def _run(cmd):
    p = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if p.returncode != 0:
        raise RuntimeError(f"command failed ({p.returncode}): {' '.join(cmd)}\n{p.stdout}")
    return p.stdout


def build(verbose: bool = False) -> str:
    """Compile every HIP source for gfx950 and link libvfxsim3d.so. Returns the library path."""
    os.makedirs(BUILD_DIR, exist_ok=True)
    os.makedirs(LIB_DIR, exist_ok=True)
    hipcc = _hipcc()
    objs = []
    cmds = []
    # every source twice: the FP32 product path and the FP64 run of the same solver (VFX_F64=1:
    # double fields, namespace vfx_f64, entry points suffixed _f64)
    for src in SOURCES:
        for f64 in (0, 1):
            obj = os.path.join(BUILD_DIR, src.replace(".hip", "_f64.o" if f64 else ".o"))
            objs.append(obj)
            cmds.append([hipcc, *HIPCC_FLAGS, f"-DVFX_F64={f64}", "-c", os.path.join(CSRC, src), "-o", obj])
    with cf.ThreadPoolExecutor(max_workers=min(8, os.cpu_count() or 1)) as ex:
        outs = list(ex.map(_run, cmds))
    if verbose:
        for c, o in zip(cmds, outs):
            print(" ".join(c))
            if o.strip():
                print(o)
    tl = _torch_lib_dir()
    tmp = LIB_PATH + ".tmp"
    # link with the host compiler against torch's unversioned runtime (DT_NEEDED libamdhip64.so),
    # not hipcc's default /opt/rocm/lib/libamdhip64.so.7
    link = ["g++", "-shared", "-fPIC", "-o", tmp, *objs, f"-L{tl}", "-lamdhip64", f"-Wl,-rpath,{tl}"]
    out = _run(link)
    if verbose and out.strip():
        print(out)
    os.replace(tmp, LIB_PATH)
    return LIB_PATH


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
