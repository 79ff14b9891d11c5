"""Build the native parts of onionpirv2_amd in-tree.

  libonionpir_gpu.so   HIP kernels + C ABI (include/onionpir_gpu.h), gfx950
  _opir_host*.so       C++ host mirror of the reference PirParams/PirClient/
                       PirServer API with its pybind11 binding

Run `python -m onionpirv2_amd.build` (or __graft_entry__.build()).
"""
import os
import subprocess
import sys
import sysconfig

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
LIBDIR = os.path.join(PKG, "lib")
GPU_LIB = os.path.join(LIBDIR, "libonionpir_gpu.so")
HOST_EXT = os.path.join(LIBDIR, "_opir_host" + sysconfig.get_config_var("EXT_SUFFIX"))

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
HIP_FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-shared",
    # packed FP32 ops stay off for every HIP translation unit
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function",
]
GPU_SRC = ["csrc/opir_kernels.hip", "csrc/opir_tree.hip", "csrc/opir_scan.hip", "csrc/opir_k2.hip",
           "csrc/opir_capi.hip", "csrc/opir_capi_k2.hip", "csrc/opir_noalloc.hip",
           "csrc/opir_seed.hip", "csrc/opir_clientgen.hip", "csrc/opir_tree_pir.hip"]
HOST_SRC = ["host/params.cpp", "host/ring.cpp", "host/prg.cpp", "host/codec.cpp", "host/client.cpp", "host/server.cpp", "host/tree.cpp",
            "host/serving.cpp", "host/bindings.cpp"]


def _run(cmd):
    print("+", " ".join(cmd), flush=True)
    subprocess.check_call(cmd, cwd=PKG)


def build_gpu():
    """Each translation unit compiles on its own (in parallel), then one link."""
    from concurrent.futures import ThreadPoolExecutor
    os.makedirs(LIBDIR, exist_ok=True)
    objdir = os.path.join(PKG, "_build")
    os.makedirs(objdir, exist_ok=True)
    cflags = [f for f in HIP_FLAGS if f != "-shared"]
    objs = [os.path.join(objdir, os.path.basename(src) + ".o") for src in GPU_SRC]
    jobs = max(1, min(len(GPU_SRC), int(os.environ.get("MAX_JOBS", os.cpu_count() or 1))))
    with ThreadPoolExecutor(jobs) as ex:
        list(ex.map(lambda so: _run([HIPCC] + cflags + ["-c", so[0], "-o", so[1]]), zip(GPU_SRC, objs)))
    tmp = GPU_LIB + ".tmp"
    _run([HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC"] + objs + ["-o", tmp])
    os.replace(tmp, GPU_LIB)


def build_host():
    os.makedirs(LIBDIR, exist_ok=True)
    import pybind11
    inc = ["-I" + pybind11.get_include(), "-I" + sysconfig.get_paths()["include"], "-I" + os.path.join(ROOT, "include")]
    tmp = HOST_EXT + ".tmp"
    _run(["g++", "-O3", "-march=x86-64-v3", "-std=c++20", "-fPIC", "-shared", "-Wall", "-Wno-unused-parameter", "-pthread"]
         + inc + HOST_SRC + ["-L" + LIBDIR, "-lonionpir_gpu", "-Wl,-rpath,$ORIGIN", "-o", tmp])
    os.replace(tmp, HOST_EXT)


def build_all():
    build_gpu()
    build_host()


if __name__ == "__main__":
    what = sys.argv[1] if len(sys.argv) > 1 else "all"
    {"gpu": build_gpu, "host": build_host, "all": build_all}[what]()
