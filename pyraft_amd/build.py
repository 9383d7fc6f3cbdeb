"""Build ``pyraft_amd/lib/libraft<EXT_SUFFIX>`` in-tree for gfx950.

    python -m pyraft_amd.build

The HIP kernels are compiled with hipcc (packed FP32 instructions disabled); the
host-side C ABI is plain C++ compiled with g++ against the HIP runtime headers
(no fused multiply-adds, so the FP64 angle tables match the reference's libm
arithmetic); hipcc links both into one shared library.
"""
import os
import subprocess
import sys
import sysconfig

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIBDIR = os.path.join(HERE, "lib")
ROCM = os.environ.get("ROCM_PATH", "/opt/rocm")
ARCH = os.environ.get("PYRAFT_AMD_ARCH", "gfx950")

HIP_SOURCES = ["ss_kernels.hip", "dda_kernels.hip", "ss_noalloc.hip"]
CXX_SOURCES = ["ss_api.cpp", "ss_batch.cpp", "dda_api.cpp", "filter_api.cpp", "io_api.cpp"]

HIPCC_FLAGS = ["-x", "hip", "--offload-arch=" + ARCH, "-O3", "-fPIC", "-std=c++17", "-fvisibility=hidden",
               "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]
CXX_FLAGS = ["-O2", "-fPIC", "-std=c++17", "-fvisibility=hidden", "-ffp-contract=off",
             "-D__HIP_PLATFORM_AMD__", "-I" + os.path.join(ROCM, "include"), "-Wall", "-Wno-unused-result"]


def library_path():
    return os.path.join(LIBDIR, "libraft" + (sysconfig.get_config_var("EXT_SUFFIX") or ".so"))


def _run(cmd):
    proc = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    out = "\n".join(l for l in proc.stdout.splitlines() if "packed-fp32-ops" not in l)
    if proc.returncode != 0:
        raise RuntimeError("command failed (%d): %s\n%s" % (proc.returncode, " ".join(cmd), out))
    return out


def build(verbose=False, out=None, defines=()):
    """Compile and link the library (into `out`, default the in-tree path), with extra
    preprocessor `defines` ("NAME=VALUE") for the HIP sources."""
    os.makedirs(BUILD, exist_ok=True)
    os.makedirs(LIBDIR, exist_ok=True)
    hipcc = os.path.join(ROCM, "bin", "hipcc")
    objs = []
    for src in HIP_SOURCES:
        obj = os.path.join(BUILD, os.path.splitext(src)[0] + ".o")
        out_txt = _run([hipcc] + HIPCC_FLAGS + ["-D" + d for d in defines] + ["-c", os.path.join(CSRC, src), "-o", obj])
        if verbose and out_txt.strip():
            print(out_txt)
        objs.append(obj)
    for src in CXX_SOURCES:
        obj = os.path.join(BUILD, os.path.splitext(src)[0] + ".o")
        out_txt = _run(["g++"] + CXX_FLAGS + ["-c", os.path.join(CSRC, src), "-o", obj])
        if verbose and out_txt.strip():
            print(out_txt)
        objs.append(obj)
    lib = out or library_path()
    tmp = lib + ".tmp"
    _run([hipcc, "--offload-arch=" + ARCH, "-shared", "-fPIC", "-o", tmp] + objs + ["-lpthread"])
    os.replace(tmp, lib)
    return lib


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
