"""Build libmcg.so (HIP, gfx950) and the pyminicore-compatible extension in-tree.

    python -m minicore_amd.build           # or: python minicore_amd/build.py

Each .hip translation unit is compiled with hipcc in parallel, then linked
into minicore_amd/libmcg.so; the pybind11 module minicore_amd/_pyminicore*.so
links against it (rpath $ORIGIN).
"""
import concurrent.futures as cf
import os
import subprocess
import sys
import sysconfig

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libmcg.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("MCG_OFFLOAD_ARCH", "gfx950")

SOURCES = ["mcg_api.hip", "mcg_dense.hip", "mcg_mstep.hip", "mcg_generic.hip", "mcg_sample.hip", "mcg_serve.hip",
           "mcg_soft.hip", "mcg_format.hip", "mcg_median.hip", "mcg_noalloc.hip", "mcg_tiles.hip", "mcg_f64.hip"]

HIP_FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-fPIC", "-std=c++17", "-Wall", "-Wno-unused-function",
    "-Wno-unused-variable", "-Wno-unused-but-set-variable",
    # packed FP32 instructions stay off for all device code
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-I", os.path.join(ROOT, "include"),
]


def _run(cmd):
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    out = "\n".join(l for l in r.stdout.splitlines() if "packed-fp32-ops" not in l)
    if r.returncode != 0:
        raise RuntimeError("command failed: " + " ".join(cmd) + "\n" + out)
    return out


def build_lib(jobs=None):
    os.makedirs(BUILD, exist_ok=True)
    objs = []
    cmds = []
    for src in SOURCES:
        obj = os.path.join(BUILD, src.replace(".hip", ".o"))
        objs.append(obj)
        cmds.append([HIPCC, *HIP_FLAGS, "-c", os.path.join(CSRC, src), "-o", obj])
    with cf.ThreadPoolExecutor(max_workers=jobs or min(len(cmds), os.cpu_count() or 4)) as ex:
        for out in ex.map(_run, cmds):
            if out.strip():
                print(out)
    tmp = LIB + ".tmp"
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *objs, "-lz"])
    # A kernel template whose body the host pass cannot instantiate loses its launch
    # stub without a compile error; the library would then fail to load.
    undef = [l.split()[-1] for l in _run(["nm", "-u", "-C", tmp]).splitlines() if "mcg::" in l]
    if undef:
        os.remove(tmp)
        raise RuntimeError("libmcg has undefined symbols of its own:\n" + "\n".join(undef[:20]))
    os.replace(tmp, LIB)
    return LIB


def build_ext():
    import pybind11
    suffix = sysconfig.get_config_var("EXT_SUFFIX")
    out = os.path.join(HERE, "_pyminicore" + suffix)
    cmd = ["g++", "-O2", "-shared", "-fPIC", "-std=c++17", "-Wall",
           "-I", pybind11.get_include(), "-I", sysconfig.get_paths()["include"],
           "-I", os.path.join(ROOT, "include"),
           os.path.join(CSRC, "pyminicore.cpp"), "-o", out + ".tmp",
           "-L", HERE, "-l:libmcg.so", "-Wl,-rpath,$ORIGIN"]
    _run(cmd)
    os.replace(out + ".tmp", out)
    return out


def build_all():
    lib = build_lib()
    ext = None
    if os.path.exists(os.path.join(CSRC, "pyminicore.cpp")):
        ext = build_ext()
    return lib, ext


if __name__ == "__main__":
    print(build_all())
    sys.exit(0)
