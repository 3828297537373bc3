"""Build every native artefact in-tree (into everest_amd/lib/).

  libtm_hip.so        hipcc, gfx950: kernels + C ABI (include/tm_abi.h)
  libeverest_core.so  g++: loader, motif reader, runtime (include/everest/data.h)
  corelib*.so         g++ + pybind11: Python module "corelib"
  libTM.so            g++: stock worker plugin (getWorker entry point)
  libCSR.so           hipcc: hand-written worker plugin reading DeviceJobData's device arrays

Run: python -m everest_amd.build   (or __graft_entry__.build()).
"""
import os
import subprocess
import sys
import sysconfig
from concurrent.futures import ThreadPoolExecutor

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
LIB = os.path.join(PKG, "lib")
INC = os.path.join(ROOT, "include")
HIP_SRC = os.path.join(PKG, "csrc", "hip")
CORE_SRC = os.path.join(PKG, "csrc", "core")

ARCH = os.environ.get("EVEREST_OFFLOAD_ARCH", "gfx950")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
CXX = os.environ.get("CXX", "g++")
# CDNA4 codegen flags: no packed-FP32 instructions anywhere in this library.
HIP_FLAGS = ["--offload-arch=" + ARCH, "-O3", "-std=c++17", "-fPIC",
             "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]
CXX_FLAGS = ["-O3", "-std=c++17", "-fPIC", "-Wall", "-Wno-unused-function"]
RPATH = ["-Wl,-rpath,$ORIGIN"]


def _run(cmd):
    p = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    out = "\n".join(l for l in p.stdout.splitlines() if "packed-fp32-ops" not in l)
    if p.returncode != 0:
        raise RuntimeError("build step failed:\n%s\n%s" % (" ".join(cmd), out))
    return out


def pybind_includes():
    import pybind11
    return [pybind11.get_include(), sysconfig.get_paths()["include"]]


def ext_suffix():
    return sysconfig.get_config_var("EXT_SUFFIX") or ".so"


HIP_UNITS = ("tm_abi.hip", "tm_sweep.hip", "tm_build.hip", "tm_server.hip", "tm_batch.hip", "tm_noalloc.hip", "tm_kv_n4_single.hip", "tm_kv_n4_multi.hip",
             "tm_kv_n4_aux.hip", "tm_kv_n8_single.hip", "tm_kv_n8_multi.hip", "tm_kv_n8_aux.hip")


def build_hip():
    """Each translation unit compiled on its own (in parallel), then linked."""
    out = os.path.join(LIB, "libtm_hip.so")
    objdir = os.path.join(LIB, "obj")
    os.makedirs(objdir, exist_ok=True)

    def unit(src):
        obj = os.path.join(objdir, os.path.splitext(src)[0] + ".o")
        _run([HIPCC] + HIP_FLAGS + ["-c", "-I" + INC, "-I" + HIP_SRC, os.path.join(HIP_SRC, src), "-o", obj])
        return obj

    with ThreadPoolExecutor(max_workers=min(len(HIP_UNITS), os.cpu_count() or 4)) as ex:
        objs = list(ex.map(unit, HIP_UNITS))
    _run([HIPCC] + HIP_FLAGS + ["-shared"] + objs + ["-pthread", "-o", out])
    return out


def build_core():
    out = os.path.join(LIB, "libeverest_core.so")
    srcs = [os.path.join(CORE_SRC, f) for f in ("loader.cc", "motif.cc", "runtime.cc", "gen.cc")]
    _run([CXX] + CXX_FLAGS + ["-shared", "-I" + INC] + srcs +
         ["-L" + LIB, "-ltm_hip", "-ldl", "-pthread"] + RPATH + ["-o", out])
    return out


def build_py():
    out = os.path.join(LIB, "corelib" + ext_suffix())
    _run([CXX] + CXX_FLAGS + ["-shared", "-fvisibility=hidden", "-I" + INC] +
         ["-I" + p for p in pybind_includes()] +
         [os.path.join(CORE_SRC, "corelib_py.cc"), "-L" + LIB, "-leverest_core", "-ltm_hip"] + RPATH + ["-o", out])
    return out


def build_plugin():
    out = os.path.join(LIB, "libTM.so")
    _run([CXX] + CXX_FLAGS + ["-shared", "-I" + INC, os.path.join(CORE_SRC, "plugin_tm.cc"),
                              "-L" + LIB, "-leverest_core", "-ltm_hip"] + RPATH + ["-o", out])
    return out


def build_csr_plugin():
    """Hand-written worker plugin bound to the reference's DeviceJobData fields."""
    out = os.path.join(LIB, "libCSR.so")
    _run([HIPCC] + HIP_FLAGS + ["-shared", "-I" + INC, os.path.join(PKG, "csrc", "plugins", "worker_csr.hip"),
                                "-L" + LIB, "-leverest_core", "-ltm_hip"] + RPATH + ["-o", out])
    return out


def build_oracle():
    subprocess.check_call(["make", "-s", "-C", os.path.join(ROOT, "oracle"), "libtm_oracle.so"])
    if os.path.isdir("/root/reference"):
        subprocess.check_call(["make", "-s", "-C", os.path.join(ROOT, "oracle"), "ref"])


def build_all(verbose=True):
    os.makedirs(LIB, exist_ok=True)
    with ThreadPoolExecutor(max_workers=2) as ex:
        f_hip = ex.submit(build_hip)
        f_or = ex.submit(build_oracle)
        f_hip.result()
        f_or.result()
    build_core()
    with ThreadPoolExecutor(max_workers=3) as ex:
        a = ex.submit(build_py)
        b = ex.submit(build_plugin)
        c = ex.submit(build_csr_plugin)
        a.result()
        b.result()
        c.result()
    if verbose:
        print("built:", ", ".join(sorted(os.listdir(LIB))))


if __name__ == "__main__":
    build_all()
    sys.exit(0)
