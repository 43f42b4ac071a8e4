"""Build the gfx950 device library and the host API (in-tree, no JIT cache).

  libcppe_hip.so   HIP kernels + the C ABI of include/cppe_hip.h      (hipcc, gfx950)
  libcppe.so       C++ host API (libcppe::CppeState, PeOptions, ...)   (g++, links libcppe_hip.so)
  pycppe*.so       pybind11 module with the reference's Python names  (g++, links both)

Run ``python -m cppe_amd.build`` (or ``__graft_entry__.build()``).
"""
import concurrent.futures as cf
import glob
import os
import subprocess
import sys
import sysconfig

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
BUILD = os.path.join(PKG, "_build")
ROCM = os.environ.get("ROCM_PATH", "/opt/rocm")
HIPCC = os.path.join(ROCM, "bin", "hipcc")

HIP_FLAGS = [
    "-O3", "-std=c++17", "-fPIC", "--offload-arch=gfx950",
    # packed FP32 instructions stay off (the compiler would otherwise form v_pk_* from float code)
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-I" + os.path.join(ROOT, "include"),
]
HOST_FLAGS = ["-O2", "-std=c++17", "-fPIC", "-pthread", "-Wall", "-Wno-unused-function",
              "-I" + os.path.join(ROOT, "include"), "-I" + os.path.join(CSRC, "host")]

LIB_HIP = os.path.join(PKG, "libcppe_hip.so")
LIB_HOST = os.path.join(PKG, "libcppe.so")


def _run(cmd):
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if r.returncode != 0:
        raise RuntimeError("build step failed:\n" + " ".join(cmd) + "\n" + r.stdout)
    return r.stdout


def _pymod_path():
    suffix = sysconfig.get_config_var("EXT_SUFFIX") or ".so"
    return os.path.join(PKG, "pycppe" + suffix)


FMM_ORDERS = range(2, 9)


def build_device(jobs=8):
    os.makedirs(BUILD, exist_ok=True)
    dev = os.path.join(CSRC, "device")
    srcs = sorted(s for s in glob.glob(os.path.join(dev, "*.hip")) if not s.endswith("fmm_order.hip"))
    steps = [([HIPCC, "-c", *HIP_FLAGS, s, "-o", os.path.join(BUILD, os.path.basename(s) + ".o")],
              os.path.join(BUILD, os.path.basename(s) + ".o")) for s in srcs]
    # the FMM kernels: one object per expansion order, compiled in parallel
    for p in FMM_ORDERS:
        o = os.path.join(BUILD, "fmm_order_p%d.o" % p)
        steps.append(([HIPCC, "-c", *HIP_FLAGS, "-DCPPE_FMM_ORDER=%d" % p, os.path.join(dev, "fmm_order.hip"),
                       "-o", o], o))
    # longest first: the octree unit pulls in hipCUB
    steps.sort(key=lambda st: 0 if "octree" in st[1] else 1)
    objs = [o for _, o in steps]
    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        list(ex.map(lambda st: _run(st[0]), steps))
    _run([HIPCC, "-shared", "-fPIC", "--offload-arch=gfx950", *objs, "-o", LIB_HIP,
          "-Wl,-rpath," + os.path.join(ROCM, "lib")])
    return LIB_HIP


def build_host(jobs=8):
    os.makedirs(BUILD, exist_ok=True)
    host_srcs = sorted(glob.glob(os.path.join(CSRC, "host", "*.cc")))
    if not host_srcs:
        return None
    objs = [os.path.join(BUILD, "host_" + os.path.basename(s) + ".o") for s in host_srcs]
    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        list(ex.map(lambda so: _run(["g++", "-c", *HOST_FLAGS, so[0], "-o", so[1]]), zip(host_srcs, objs)))
    _run(["g++", "-shared", "-fPIC", "-pthread", *objs, "-o", LIB_HOST, "-L" + PKG, "-lcppe_hip",
          "-Wl,-rpath,$ORIGIN"])
    py_srcs = sorted(glob.glob(os.path.join(CSRC, "python", "*.cc")))
    if py_srcs:
        import pybind11
        inc = [pybind11.get_include(), sysconfig.get_paths()["include"]]
        _run(["g++", "-shared", *HOST_FLAGS, *["-I" + i for i in inc], *py_srcs, "-o", _pymod_path(),
              "-L" + PKG, "-lcppe", "-lcppe_hip", "-Wl,-rpath,$ORIGIN"])
    return LIB_HOST


def build(jobs=8):
    build_device(jobs)
    build_host(jobs)


def built():
    return os.path.exists(LIB_HIP)


if __name__ == "__main__":
    build()
    print("built", LIB_HIP, file=sys.stderr)
