"""In-tree build of the native parts of keldy_amd.

libkeldy_amd.so   HIP kernels (gfx950) + the C-ABI of include/keldy_amd.h (hipcc)
_keldy_host*.so   C++ host API (model, warpers, Sobol, integrator) + pybind11
                  bindings; links libkeldy_amd.so through an $ORIGIN rpath
"""
import os
import subprocess
import sys
import sysconfig
from concurrent.futures import ThreadPoolExecutor

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
LIBDIR = os.path.join(PKG, "lib")
BUILD = os.path.join(ROOT, "_build")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"

HIP_FLAGS = ["-O3", f"--offload-arch={ARCH}", "-fPIC", "-std=c++17",
             # keep packed FP32 (v_pk_*_f32) out of the generated code
             "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]

HIP_SOURCES = [
    "kernels/kd_integrand.hip",
    "kernels/kd_tree.hip",
    "kernels/kd_flow.hip",
    "kernels/kd_cofactor.hip",
    "kernels/kd_split.hip",
    "kernels/kd_wsplit.hip",
    "kernels/kd_ctree.hip",
    "capi/kd_capi.cpp",
]

HOST_SOURCES = [
    "host/model.cpp",
    "host/quadrature.cpp",
    "host/warpers.cpp",
    "host/qrng.cpp",
    "host/integrator.cpp",
    "bindings/py_keldy.cpp",
]

LIB_NAME = "libkeldy_amd.so"


def lib_path():
    return os.path.join(LIBDIR, LIB_NAME)


def host_ext_path():
    suffix = sysconfig.get_config_var("EXT_SUFFIX") or ".so"
    return os.path.join(PKG, "_keldy_host" + suffix)


def _run(cmd):
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("build step failed: " + " ".join(cmd[:3]) + " ...")
    return r


def _obj(src):
    base = src.replace("/", "_").rsplit(".", 1)[0]
    return os.path.join(BUILD, base + ".o")


def build_device_lib():
    os.makedirs(BUILD, exist_ok=True)
    os.makedirs(LIBDIR, exist_ok=True)
    srcs = [s for s in HIP_SOURCES if os.path.exists(os.path.join(CSRC, s))]

    def compile_one(src):
        cmd = [HIPCC, *HIP_FLAGS, "-c", os.path.join(CSRC, src), "-o", _obj(src)]
        if src.endswith(".cpp"):
            cmd[1:1] = ["-x", "hip"]
        _run(cmd)

    with ThreadPoolExecutor(max_workers=min(4, len(srcs))) as ex:
        list(ex.map(compile_one, srcs))
    tmp = lib_path() + ".tmp"
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-o", tmp, *[_obj(s) for s in srcs]])
    os.replace(tmp, lib_path())
    return lib_path()


def build_host_ext():
    import pybind11

    srcs = [os.path.join(CSRC, s) for s in HOST_SOURCES if os.path.exists(os.path.join(CSRC, s))]
    inc = [f"-I{pybind11.get_include()}", f"-I{sysconfig.get_paths()['include']}",
           f"-I{os.path.join(ROOT, 'include')}", f"-I{CSRC}"]
    out = host_ext_path()
    tmp = out + ".tmp"
    cmd = ["g++", "-O2", "-std=c++17", "-fPIC", "-shared", "-fopenmp", "-fvisibility=hidden", *inc, *srcs,
           f"-L{LIBDIR}", "-lkeldy_amd", "-Wl,-rpath,$ORIGIN/lib", "-o", tmp]
    _run(cmd)
    os.replace(tmp, out)
    return out


def build_all():
    build_device_lib()
    build_host_ext()


if __name__ == "__main__":
    what = sys.argv[1] if len(sys.argv) > 1 else "all"
    {"device": build_device_lib, "host": build_host_ext, "all": build_all}[what]()
