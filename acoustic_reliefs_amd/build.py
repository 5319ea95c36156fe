"""Builds libarh.so (HIP, gfx950) and the acoustics3d_impl pybind11 extension in-tree.

    python -m acoustic_reliefs_amd.build          # or __graft_entry__.build()

Every HIP translation unit is compiled with hipcc --offload-arch=gfx950 -O3 and with packed
FP32 instructions disabled (-Xclang -target-feature -Xclang -packed-fp32-ops). The extension
module is plain C++ (g++) and reaches the GPU only through the C-ABI in include/arh.h.
"""
import os
import subprocess
import sys
import sysconfig
from concurrent.futures import ThreadPoolExecutor

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
BUILD = os.path.join(ROOT, "_build")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"

HIP_SOURCES = ["geometry.hip", "assemble.hip", "matvec.hip", "matvec64.hip", "matvec_mfma.hip", "solver.hip", "receivers.hip",
               "grad.hip", "qaca.hip", "relief.hip", "capi.cpp", "partition.cpp", "serve.cpp", "pool.cpp"]
HIP_FLAGS = ["--offload-arch=" + ARCH, "-O3", "-fPIC", "-std=c++17", "-Xclang", "-target-feature", "-Xclang",
             "-packed-fp32-ops", "-Wall", "-Wno-unused-function", "-Wno-unused-result", "-I" + os.path.join(ROOT, "include")]


def _run(cmd):
    p = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if p.returncode != 0:
        raise RuntimeError("build failed: %s\n%s" % (" ".join(cmd), p.stdout))
    return p.stdout


def ext_suffix():
    return sysconfig.get_config_var("EXT_SUFFIX") or ".so"


def lib_path():
    return os.path.join(PKG, "libarh.so")


def module_path():
    return os.path.join(PKG, "acoustics3d_impl" + ext_suffix())


def build_lib(jobs=8):
    os.makedirs(BUILD, exist_ok=True)
    srcs = [s for s in HIP_SOURCES if os.path.exists(os.path.join(CSRC, s))]

    def compile_one(src):
        obj = os.path.join(BUILD, src.replace(".", "_") + ".o")
        lang = ["-x", "hip"] if src.endswith(".hip") else []
        _run([HIPCC] + HIP_FLAGS + lang + ["-c", os.path.join(CSRC, src), "-o", obj])
        return obj

    with ThreadPoolExecutor(max_workers=jobs) as ex:
        objs = list(ex.map(compile_one, srcs))
    tmp = lib_path() + ".tmp"
    _run([HIPCC, "--offload-arch=" + ARCH, "-shared", "-fPIC", "-o", tmp] + objs)
    os.replace(tmp, lib_path())
    return lib_path()


def build_module():
    import pybind11
    inc = [sysconfig.get_paths()["include"], pybind11.get_include(), os.path.join(ROOT, "include")]
    tmp = module_path() + ".tmp"
    cmd = ["g++", "-O2", "-shared", "-fPIC", "-std=c++17", "-fvisibility=hidden", "-Wall"]
    cmd += ["-I" + i for i in inc]
    cmd += [os.path.join(CSRC, "module.cpp"), "-L" + PKG, "-larh", "-Wl,-rpath,$ORIGIN", "-o", tmp]
    _run(cmd)
    os.replace(tmp, module_path())
    return module_path()


def build(jobs=8):
    build_lib(jobs)
    build_module()


if __name__ == "__main__":
    build()
    print("built", lib_path(), module_path())
    sys.exit(0)
