"""Build libtopaz_amd.so (HIP kernels + C ABI) in-tree for gfx950.

    python -m topaz_amd.build        # or topaz_amd.build.build()
"""
import os
import shutil
import subprocess
import sys
from pathlib import Path

PKG = Path(__file__).resolve().parent
CSRC = PKG / "csrc"
LIB = PKG / "libtopaz_amd.so"
SOURCES = ["runtime.hip", "sw_kernels.hip", "gapless_kernels.hip", "traceback_kernels.hip", "compat.hip",
           "server.hip", "suffix_array.hip", "gapless_noalloc.hip", "walk.hip", "search.hip", "sw_pair_kernels.hip",
           "walk_kernels.hip", "sw_lane_kernels.hip", "seg.hip", "compat_scalar.hip"]
ARCH = os.environ.get("TOPAZ_AMD_ARCH", "gfx950")


def hipcc():
    for c in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if c and Path(c).exists():
            return c
    raise RuntimeError("hipcc not found")


def build(verbose=False):
    objs = []
    build_dir = PKG / "_build"
    build_dir.mkdir(exist_ok=True)
    flags = ["--offload-arch=" + ARCH, "-O3", "-std=c++17", "-fPIC",
             "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
             "-Wall", "-Wno-unused-function"]
    procs = []
    for src in SOURCES:
        obj = build_dir / (Path(src).stem + ".o")
        cmd = [hipcc()] + flags + ["-c", str(CSRC / src), "-o", str(obj)]
        procs.append((cmd, subprocess.Popen(cmd, stdout=subprocess.PIPE,
                                            stderr=subprocess.STDOUT, text=True)))
        objs.append(obj)
    for cmd, p in procs:
        out, _ = p.communicate()
        out = "\n".join(l for l in out.splitlines()
                        if "packed-fp32-ops' is not a recognized feature" not in l)
        if p.returncode != 0:
            raise RuntimeError("compile failed: %s\n%s" % (" ".join(cmd), out))
        if verbose and out.strip():
            print(out)
    tmp = LIB.with_suffix(".so.tmp")
    cmd = [hipcc(), "--offload-arch=" + ARCH, "-shared", "-fPIC", "-o", str(tmp)] + \
        [str(o) for o in objs]
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if r.returncode != 0:
        raise RuntimeError("link failed: %s\n%s" % (" ".join(cmd), r.stdout))
    os.replace(tmp, LIB)
    # the gapless leg's calibration program (random window reads, no compute)
    probe = PKG / "probe_window_gather"
    cmd = [hipcc()] + [f for f in flags if f not in ("-fPIC",)] + \
        [str(CSRC / "probe_window_gather.hip"), "-o", str(probe) + ".tmp"]
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if r.returncode != 0:
        raise RuntimeError("compile failed: %s\n%s" % (" ".join(cmd), r.stdout))
    os.replace(str(probe) + ".tmp", probe)
    return LIB


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
